// cd_layout.h -- memory layout of the codon-design engine's device state,
// shared by the kernels and the host API.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace cd {

constexpr int NOBJ = 6;                  // common.cuh:23
constexpr int NCODON = 64;               // common.cuh:17
constexpr int NSENSE = 61;               // TOTAL_CODON_NUM - STOP_CODON_NUM
constexpr int STOP_AA = 20;              // 'Z'
constexpr int MIN_HAIRPIN = 4;           // common.cuh:34
constexpr float F_PRECISION = 0.000001f; // common.cuh:53
constexpr int GC_UP = 1, GC_DOWN = -1;   // common.cuh:42-43
constexpr int MAX_M = 16;                // CDS copies per design supported
constexpr int MAX_SYN = 6;               // most synonyms of one amino acid (Leu, Ser, Arg)
constexpr int MAX_W = 256;               // dominance words per row: pools up to 16384

// Philox stream tags: counter = (generation, design | child | ref, slot, tag)
constexpr uint32_t TAG_INIT = 0x494E4954u;
constexpr uint32_t TAG_TYPE = 0x54595045u;
constexpr uint32_t TAG_MUT = 0x4D555441u;
constexpr uint32_t TAG_SL = 0x534C5452u;
constexpr uint32_t TAG_NICHE = 0x4E494348u;

// Per-CDS statistics kept with every design so that a child recomputes only
// the CDS its operator changed.
struct CdsStat {
    int64_t cai;  // sum of 2^40-scaled log weights
    int64_t cpb;  // sum of 2^40-scaled pair scores over pairs k,k+1, k < L-2
    int32_t hsc;  // hidden stop codons over junctions k|k+1, k < L-1
    int32_t gc3;  // codons whose third base is G or C
    int32_t l, p, q;  // longest stem and its position (canonical tie rule)
    int32_t pad;      // 1 while the stem fields await k_stem
    uint64_t key;     // best stem key of the parts of k_stem (atomic max), valid while pad is 1
    float ov[4];      // the CDS's CAI, CPB, HSC quotient and GC3 fraction, from the sums above
};
static_assert(sizeof(CdsStat) == 64, "CdsStat records move as four 16-byte words");

// Selection state of one job (the reference keeps it in __device__ globals,
// common.cuh:46-59).
struct SelState {
    uint32_t gen;      // generations completed
    int32_t n_cut;     // fronts filled N exactly (no niching)
    int32_t hyp;       // hyperplane exception (MNDF nadir)
    int32_t ranked;    // designs ranked
    int32_t last;      // index of the last (cut) front
    int32_t base;      // designs in fronts before the last
    uint32_t pad;
    unsigned long long cds_scanned;  // CDS whose objectives were recomputed (cumulative)
    float ext[NOBJ * NOBJ];
    float ideal[NOBJ], worst[NOBJ], nadir[NOBJ];
};

// Per-generation accumulators of the selection, filled with atomics by the
// workgroups of k_dominance, k_front0 and k_levels and cleared with the
// dominator counts before every sort (acc_clear_word).  Float extrema are kept
// as order-preserving bits (ord_bits) so that integer atomics combine them.
struct SelAcc {
    uint32_t lo[NOBJ];    // smallest value of each objective over the pool
    uint32_t hi[NOBJ];    // largest value over the pool
    uint32_t f0hi[NOBJ];  // largest value over the first front
    int32_t f0;           // size of the first front
    int32_t ncut;         // designs of the cut buckets (those ranked by priority)
    unsigned long long asf[NOBJ];  // per axis: (ASF bits << 32 | design) minimum over the first front
};
constexpr int ACC_WORDS = (int)(sizeof(SelAcc) / 4);
static_assert(sizeof(SelAcc) % 8 == 0, "SelAcc is cleared in 32-bit words and holds u64 minima");

// Order-preserving bits of a float: unsigned comparison of ord_bits(a) and
// ord_bits(b) orders a and b as floats (-0 below +0; not for NaN).
__host__ __device__ inline uint32_t ord_bits(float f) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float ord_float(uint32_t u) {
    return __builtin_bit_cast(float, (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}
// The value word i of a cleared SelAcc holds: minima start at all ones,
// everything else at zero.
__host__ __device__ inline uint32_t acc_clear_word(int i) {
    const int lo0 = (int)(offsetof(SelAcc, lo) / 4), asf0 = (int)(offsetof(SelAcc, asf) / 4);
    return ((i >= lo0 && i < lo0 + NOBJ) || i >= asf0) ? 0xFFFFFFFFu : 0u;
}

// Per-reference-point arrays of the niching (DevJob::niche, R entries each):
// int2 (designs of the earlier fronts, designs of the last front); int picks,
// level-0 pick, offset of a cut bucket in the cut list (-1 when not cut), its
// fill count; u64 (distance bits << 32 | design) minimum over the bucket
// (its nearest design).
struct NicheView {
    int2 *cr;
    int32_t *kp, *lvl0, *coff, *fill;
    unsigned long long *nearest;
};
constexpr int NICHE_INTS = 8;  // 32-bit words per reference point
__host__ __device__ inline NicheView niche_view(int32_t *base, int R) {
    NicheView v;
    v.cr = (int2 *)base;
    v.kp = base + 2 * R;
    v.lvl0 = base + 3 * R;
    v.coff = base + 4 * R;
    v.fill = base + 5 * R;
    v.nearest = (unsigned long long *)(base + 6 * R);
    return v;
}
constexpr int NICHE_STAGE = 1024;  // reference points whose niche arrays fit in a workgroup's LDS

// Device view of one job.  Pools ping-pong on generation parity: generation
// g reads buffer g&1 and writes buffer (g+1)&1.
// Stride of a CDS row of L codons in the genome pools and in k_design's LDS:
// L rounded up to 16 bytes, so rows move in 16-byte words.
__host__ __device__ constexpr int row_stride(int L) { return (L + 15) & ~15; }

// A codon's code is its three nucleotides A=0 C=1 G=2 U=3 in six bits,
// first << 4 | second << 2 | third; the 64 codons have the 64 codes.
struct DevJob {
    int32_t L, m, N, P, R, W, den_sl;
    int32_t Ls;            // stride of a CDS row in the genome pools: L rounded up to 16 bytes
    uint64_t seed, thr_pm, thr_2pm;
    uint64_t stop6;        // bit c set when code c is a stop codon (the table's last three codons)
    float gc3_ref, pad1;
    const uint8_t *pos;    // [2][Ls]: per position the first codon of its amino acid, then its synonym count
                           // (padding positions: codon 0, one synonym)
    const uint8_t *code;   // [64] code of each codon of the table
    const int64_t *lw6;    // [64] 2^40-scaled log weights by code
    const int64_t *cpfx6;  // [64*64] 2^40-scaled pair scores by codes (0 where either is a stop)
    const float *cps6;     // [64*64] pair scores by codes (0 where either is a stop)
    const float *refpts;   // [R*6]
    int8_t *genome[2];     // [P][m][Ls] synonym indices, CDS rows padded with zeros to Ls
    float *obj[2];         // [P][6]
    int8_t *oidx[2];       // [P][12]
    int32_t *pql[2];       // [P][4] (p, q, l, 0)
    CdsStat *cst[2];       // [P][m]
    int32_t *hdp[2];       // [P][m(m-1)/2] Hamming distances of the design's CDS pairs (i < j, lexicographic)
    uint2 *planes;         // [P][m][2][scan_words(3L)] bit planes (sequence, reverse complement) of the CDS
                           // k_design evaluated, read by k_stem
    int32_t *front_of;     // [P]
    int32_t *np;           // [P] dominator counts
    int32_t *stem_max;     // [P] longest stem of the design (k_design: unchanged CDS, k_stem: atomic maximum)
    uint64_t *dom;         // [W][P], word-major: bit b of dom[w * P + j] set when design 64w + b dominates j
    int32_t *ref_of;       // [P]
    float *dist;           // [P]
    int32_t *niche;        // NICHE_INTS x R words per reference point (NicheView); null for a sort alone
    SelState *st;
    SelAcc *acc;
    uint64_t *ranked;      // [W] first front (bit i of word w: design 64w + i)
    uint64_t *selm;        // [W] survivors: the next generation's parents, in ascending pool index; null for a sort alone
    uint32_t *prio;        // [P] high word of a last-front design's niche priority
    uint64_t *cutl;        // [P] (priority << 32 | design) of the cut buckets' designs, bucket by bucket
    int32_t *cutr;         // [P] the reference point of each cut-list entry
};

}  // namespace cd
