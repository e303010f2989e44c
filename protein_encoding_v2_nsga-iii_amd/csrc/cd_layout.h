// cd_layout.h -- memory layout of the codon-design engine's device state,
// shared by the kernels and the host API.
#pragma once

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
constexpr uint32_t TAG_MUTB = 0x4D555442u;
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
};

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

// Dynamic LDS of k_niche: the bucketed last front (design indices and
// priority high words, 8P bytes) always; with tier >= 1 also the bucket's
// distances (4P) and the per-reference-point arrays (six int arrays and one
// u64 array, 32R); with tier 2 also copies of front_of, ref_of and the
// survivor flags (9P).  What is not in LDS lives in the job's global scratch.
constexpr size_t NICHE_LDS_CAP = 150 * 1024;

// P-sized arrays are laid out with an even stride so that the u64 array that
// follows them stays 8-byte aligned.
__host__ __device__ inline int niche_stride(int P) { return (P + 1) & ~1; }

__host__ __device__ inline int niche_tier(int P, int R) {
    const size_t Pa = niche_stride(P);
    if (21 * Pa + (size_t)32 * R + 16 <= NICHE_LDS_CAP) return 2;
    if (12 * Pa + (size_t)32 * R <= NICHE_LDS_CAP) return 1;
    return 0;
}

__host__ __device__ inline size_t niche_lds_need(int P, int R) {
    const int t = niche_tier(P, R);
    const size_t Pa = niche_stride(P);
    return 8 * Pa + (t >= 1 ? 4 * Pa + (size_t)32 * R : 0) + (t >= 2 ? 9 * Pa + 16 : 0);
}

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
    const int8_t *aa;      // [L]
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
    uint2 *planes;         // [P][m][2][scan_words(3L)] bit planes (sequence, reverse complement) of the CDS
                           // k_design evaluated, read by k_stem
    int32_t *sorted;       // [N]
    int32_t *front_of;     // [P]
    int32_t *np;           // [P] dominator counts
    int32_t *stem_max;     // [P] longest stem of the design (k_design: unchanged CDS, k_stem: atomic maximum)
    uint64_t *dom;         // [P][W]: bit i of row j set when i dominates j
    int32_t *ref_of;       // [P]
    float *dist;           // [P]
    int32_t *niche;        // [8][R]: count, rem, off, fill, kpick, level-0 pick, then u64 [R] pick threshold
    uint8_t *flag;         // [P]
    SelState *st;
};

}  // namespace cd
