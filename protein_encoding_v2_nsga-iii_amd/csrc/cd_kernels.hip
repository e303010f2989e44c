// cd_kernels.hip -- HIP kernels of the codon-design engine for gfx950.
//
// One generation is these launches over every resident job:
//   k_design<OFFSPRING>  one workgroup per child: parent from the survivor
//                        mask, survivor copy, mutation, objectives 0-4
//                        (mutationKernel, main.cu:121-240)
//   k_stem_list          longest stem of each CDS the mutation changed
//                        (calMaximumSL)
//   k_dominance          objective 5 of the children, bit-packed
//                        eps-domination matrix (sorting.cuh:994-1023), the
//                        pool's ideal and worst points
//   k_front0             the first front, its maxima and ASF minima, many
//                        workgroups per job
//   k_fronts             one workgroup per job: further fronts, extreme
//                        points, hyperplane/MNDF nadir (sorting.cuh:974-1089,
//                        132-972)
//   k_associate          nearest reference line per ranked design and the
//                        niche counts (sorting.cuh:1105-1180)
//   [k_levels]           the niche levels, when the reference points do not
//                        fit in LDS (else k_place settles them)
//   k_place              level-0 picks, whole buckets, the cut list
//   k_niche              the cut buckets ranked by priority
//                        (sorting.cuh:1182-1279); ends the generation
// Hand-offs between workgroups happen at kernel boundaries; atomics that
// fold per-job quantities are reduced over the workgroup first (same-address
// atomics serialise at the device's coherence point).
// Work lists map blockIdx.x to (job, item) so the same kernels advance one
// job or a batch of uneven jobs.
#include <cfloat>
#include <mutex>

#include "cd_device.h"
#include "cd_kernels.h"
#include "cd_qgram.h"
#include "cd_scan.h"
#include "cd_tile.h"

namespace cd {

__host__ __device__ constexpr int align16(int x) { return (x + 15) & ~15; }

// k_design's dynamic LDS: log weights by codon code lw6 [64] (int64), the
// code of every codon of the table [64], per position the first codon and
// the number of synonyms of its amino acid [Ls] each, then the design's CDS
// rows (stride Ls = row_stride(L)): synonym indices syn [m][Ls] and codon
// codes cd6 [m][Ls], then the CDS statistics [m].  Rows are 16-byte aligned,
// so the genome moves between HBM and LDS in 16-byte words and four codes
// share a 32-bit word in the Hamming sums.
// After the genome row is stored, the syn area stages the forward bit planes
// of the changed CDS (plane_stage_bytes each), so its stride per CDS is the
// larger of the two.
constexpr int DL_LW6 = 0, DL_CODE = 512, DL_POS = 576;
static_assert(DL_POS % 16 == 0, "design LDS rows are 16-byte aligned");

// forward planes of one CDS staged in LDS: scan_words(3L) words and one zero
// word past them (the reverse-complement windows read one word ahead)
__host__ __device__ inline int plane_stage_bytes(int L) { return align16((scan_words(3 * L) + 1) * 8); }
__host__ __device__ inline int syn_stride(int L) {
    const int a = row_stride(L), b = plane_stage_bytes(L);
    return a > b ? a : b;
}

size_t design_lds_bytes(int L, int m) {
    const size_t Ls = row_stride(L), SA = syn_stride(L);
    return (size_t)DL_POS + 2 * Ls + (size_t)m * (SA + Ls) + align16(m * (int)sizeof(CdsStat)) + 16;
}

// one CDS's bit planes: the sequence and its reverse complement, each
// scan_words(3L) words of (low bits, high bits) of the nucleotide codes
size_t cds_plane_bytes(int L) { return (size_t)2 * scan_words(3 * L) * sizeof(uint2); }

// dynamic LDS of k_stem: the bit planes of one CDS, then the q-gram search's
// index and hit queue
size_t stem_lds_bytes(int L) { return cds_plane_bytes(L) + qg_lds_bytes(3 * L); }

constexpr int HITQ = 192;  // a wave's queue of hits: a full batch of 64 and one step's 128
// k_design's static LDS for a workgroup of NW waves.
template <int NW>
struct DesignShared {
    float pobj[NOBJ];
    int32_t pidx[2 * NOBJ];
    int32_t ppql[4];
    int32_t dirty[MAX_M];
    int32_t hd[MAX_M * (MAX_M - 1) / 2];
    int64_t red[8 * NW];
    int32_t red32[8 * NW];
    uint64_t redu[16];
    int32_t misc[16];
    int32_t slw[NW * 11];  // mut_sl: per wave, first [4], last [4], full [3]
    int32_t dlist[MAX_M];  // the changed CDS, in order
    uint16_t hitq[NW][HITQ];  // per wave: codons that passed an operator's test, awaiting their work
};

// CdsStat records in global memory, four 16-byte words through the global
// address space (an aggregate copy would go through flat accesses).
struct CdsWords {
    u32x4 w[4];
};
static_assert(sizeof(CdsWords) == sizeof(CdsStat), "a CdsStat record is four 16-byte words");
__device__ __forceinline__ CdsStat gload_stat(const CdsStat *p) {
    const u32x4 *q = (const u32x4 *)p;
    CdsWords v;
#pragma unroll
    for (int i = 0; i < 4; ++i) v.w[i] = gload(q + i);
    return __builtin_bit_cast(CdsStat, v);
}
__device__ __forceinline__ void gstore_stat(CdsStat *p, const CdsStat &s) {
    u32x4 *q = (u32x4 *)p;
    const CdsWords v = __builtin_bit_cast(CdsWords, s);
#pragma unroll
    for (int i = 0; i < 4; ++i) gstore(q + i, v.w[i]);
}

// k_design's LDS tables and the per-position synonym ranges.
struct DesignLds {
    const int64_t *lw6;    // [64] log weights by code
    const uint8_t *code;   // [64] code of each codon of the table
    const uint8_t *pbase;  // [Ls] first codon of position k's amino acid
    const uint8_t *pnum;   // [Ls] synonyms of position k's amino acid
    uint64_t stop6;        // stop codes
    int Ls;                // row stride
};

// Facts of codons given by their codes (first base << 4 | second << 2 | third).
// Third base C (1) or G (2).
__device__ __forceinline__ int code_gc3(int a) { return (a ^ (a >> 1)) & 1; }
// Hidden stop codons read across the junction a|b: a[1] a[2] b[0] and
// a[2] b[0] b[1] (countRightSideHSC, utils.cuh:138-161).
__device__ __forceinline__ int code_hidden(uint64_t stop6, int a, int b) {
    return (int)((stop6 >> (((a & 15) << 2) | (b >> 4))) & 1u) + (int)((stop6 >> (((a & 3) << 4) | (b >> 2))) & 1u);
}
// Nucleotide mismatches of the codons whose codes are the bytes of x ^ y:
// a byte's three 2-bit fields differ when either bit of the field is set.
__device__ __forceinline__ int code_mismatch4(uint32_t x) { return __popc((x | (x >> 1)) & 0x15151515u); }
__device__ __forceinline__ int code_mismatch(int a, int b) { return code_mismatch4((uint32_t)(a ^ b)); }
// Base p of a CDS given as a row of codes.
__device__ __forceinline__ int code_base(const uint8_t *row, int p) {
    const unsigned k = (unsigned)p / 3u;
    return (row[k] >> (4 - 2 * (p - 3 * (int)k))) & 3;
}
// ---------------------------------------------------------------------------
// Longest stem of one CDS held as bit planes in LDS (calMaximumSL,
// utils.cuh:712-888; the planes are built by k_design).  A changed CDS of an
// offspring generation is searched by q-grams (cd_qgram.h) from its parent's
// stem; the initial population, and a CDS the search cannot decide, by tiles
// (cd_tile.h): the lanes of the `parts` workgroups of the CDS split its tiles
// into contiguous ranges.  Each workgroup takes the maximum key (longest, then
// smallest p, then smallest q).  A CDS with no run of 4 and no seed of 4 is
// scanned exactly along its anti-diagonals (cd_scan.h).
// ---------------------------------------------------------------------------
// The seed (wave 0): the longest run of pairs (i - k, j + k), k < 64, that
// still pair in the child, with (i, j) = (p + l - 1, q) the parent's
// innermost pair.  Each such run is a stem of the child (its pairs lie on one
// anti-diagonal and enclose at least the parent's loop), so its length bounds
// the longest stem from below.
__device__ __forceinline__ int plane_base(const uint2 *sp, int p) {
    const uint2 w = sp[p >> 5];
    const int b = p & 31;
    return (int)(((w.x >> b) & 1u) | (((w.y >> b) & 1u) << 1));
}
__device__ __forceinline__ int stem_seed(const uint2 *sp, int n, int l, int p, int q, int lane) {
    const int i = p + l - 1, j = q;
    const bool pair = l > 0 && lane <= i && j + lane < n && ((plane_base(sp, i - lane) ^ plane_base(sp, j + lane)) == 3);
    uint64_t m = __ballot(pair);
    int run = 0;
    while (m) {  // each round removes the last pair of every run
        ++run;
        m &= m >> 1;
    }
    return run;
}

// q-gram search (cd_qgram.h) of a CDS for the runs of at least K.  The
// workgroup builds the index of s in LDS (head, then the links); lane t walks
// the lookups of the chunks t, t + NT, ... of 8 rc positions.  In rounds,
// each wave walks until its segment of the hit queue could overflow or its
// lanes are done, appending the hits of each step by ballot, and then the
// whole workgroup measures the queued hits.  ctl[4..15]: per wave, the hits
// queued, whether lanes have lookups left, whether one ran out of budget.
// Sets *best to the workgroup's maximum key and returns whether every lookup
// finished.
__device__ __forceinline__ bool stem_qgram(int n, const uint2 *sp, const uint2 *rp, int K, uint16_t *head, int32_t *ctl,
                           uint64_t *redu, uint64_t &best) {
    const int tid = threadIdx.x, NT = blockDim.x, lane = tid & 63, wid = tid >> 6, nw = NT >> 6;
    const int st = qg_stride(K), NI = qg_index_count(n, st);
    uint16_t *next = head + QG_CODES;
    uint32_t *queue = (uint32_t *)((uint8_t *)next + qg_link_bytes(n));
    uint4 *h4 = (uint4 *)head;
    // two empty heads (QG_NIL) per word; made here rather than kept in
    // registers across the search loop and the tile fallback's call
    uint32_t nil2;
    asm volatile("v_mov_b32 %0, -1" : "=v"(nil2));
    __syncthreads();  // an earlier search's index and queue are no longer read
    for (int i = tid; i < QG_CODES / 8; i += NT) h4[i] = make_uint4(nil2, nil2, nil2, nil2);
    __syncthreads();
    for (int k = tid; k < NI; k += NT) qg_insert(sp, st, k, head, next);
    __syncthreads();
    const int xe = n - QG_Q + 1, nch = (xe + 7) >> 3;
    QgLane ln;
    qg_lane_start(ln, rp, head, tid, nch, xe);
    int budget = qg_budget(8 * ((nch + NT - 1) / NT));
    uint32_t *seg = queue + wid * QG_SEG;
    uint64_t b = 0;
    bool over = false;
    for (;;) {
        int cnt = 0;  // the same in every lane of the wave
        while (cnt <= QG_SEG - 64) {
            const bool live = !qg_lane_done(ln, nch) && budget > 0 && !qg_chunk_done(ln);
            if (!__any(live)) {
                // every lane of the wave has walked its chunk: the lanes with
                // chunks left load their next ones together (a wave's steps
                // then only read chain entries)
                const bool adv = !qg_lane_done(ln, nch) && budget > 0;
                if (!__any(adv)) break;
                if (adv) qg_lane_next(ln, rp, head, nch, NT, xe);
                continue;
            }
            int a = 0, x = 0;
            const bool hit = live && qg_walk_step(ln, next, n, st, budget, a, x);
            const uint64_t bal = __ballot(hit);
            const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            if (hit) seg[cnt + rank] = qg_pack(a, x);
            cnt += __popcll(bal);
        }
        const bool left = !qg_lane_done(ln, nch);
        const bool rem = __any(left), out = __any(left && budget <= 0);
        if (lane == 0) {
            ctl[4 + wid] = cnt;
            ctl[8 + wid] = rem;
            ctl[12 + wid] = out;
        }
        __syncthreads();
        int total = 0, more = 0;
        for (int w = 0; w < nw; ++w) {
            total += ctl[4 + w];
            more |= ctl[8 + w];
            over |= ctl[12 + w] != 0;
        }
        for (int h = tid; h < total; h += NT) {  // hit h: wave w's entry h minus the counts before it
            int w = 0, r = h;
            while (r >= ctl[4 + w]) r -= ctl[4 + w++];
            const uint32_t v = queue[w * QG_SEG + r];
            qg_measure(sp, rp, n, (int)(v >> 16), (int)(v & 0xFFFFu), K, b);
        }
        __syncthreads();  // the counts and the queue are read before the next round writes them
        if (!more || over) break;
    }
    best = block_max_u64(b, redu);
    return !over;
}

// The tiled scan of share `part` of `parts` of a CDS with a stem of `seed`
// known (0: none), and its exact scan when there is neither a seed of 4 nor
// a run of 4; returns the workgroup's maximum key.
__device__ __forceinline__ uint64_t stem_tiles(int n, const uint2 *sp, const uint2 *rp, int seed, int part, int parts,
                                               uint64_t *redu) {
    const int tid = threadIdx.x, NT = blockDim.x;
    // Every lane's K is a length some stem of this CDS has, so the wave may
    // share the largest after each step.  Trip counts are the same for every
    // lane, so every lane takes part in the exchanges.
    int K = max(seed, TILE_MIN_RUN);
    uint64_t best = 0;
    {
        const int TT = 2 * tile_prefix(n, tile_blocks(n)), T = parts * NT, L = part * NT + tid;
        const int t0 = (int)((int64_t)L * TT / T), t1 = (int)((int64_t)(L + 1) * TT / T);
        const int trips = (TT + T - 1) / T;
        TileCursor cur;
        for (int j = 0; j < trips; ++j) {
            if (t0 + j < t1) {
                if (j == 0) tile_seek(sp, rp, n, t0, cur);
                tile_step(sp, rp, n, cur, K, best);
            }
            K = wave_max_nonneg(K);
        }
    }
    best = block_max_u64(best, redu);
    // With a seed of 4 or more the longest stem is at least 4 long, so the
    // share holding it measures it; a share without a flag holds nothing
    // longer.  Without one, a share with no run of 4 is scanned exactly.
    if (seed < TILE_MIN_RUN && (int)(best >> 40) < TILE_MIN_RUN) {
        best = 0;
        K = 1;
        const int stride = parts * NT, S0 = MIN_HAIRPIN + 1 + part * NT + tid;
        const int Sw = MIN_HAIRPIN + 1 + part * NT + (tid & ~63), last = 2 * n - 3;
        const int trips = Sw <= last ? (last - Sw) / stride + 1 : 0;
        for (int j = 0; j < trips; ++j) {
            const int S = S0 + j * stride;
            if (S <= last) scan_diagonal(sp, rp, n, S, MIN_HAIRPIN, K, best);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) K = max(K, __shfl_xor(K, o, 64));
        }
        best = block_max_u64(best, redu);
    }
    return best;
}

// The whole CDS's tiled scan out of line, for the offspring search's rare
// fallback: inlined, its registers would raise the pressure of the search
// path.  The LDS pointers keep their address space, so the scan's loads stay
// LDS loads.
typedef __attribute__((address_space(3))) uint2 lds_uint2;
typedef __attribute__((address_space(3))) uint64_t lds_uint64;
__device__ __noinline__ uint64_t stem_tiles_whole(int n, const lds_uint2 *planes, int seed, lds_uint64 *redu) {
    const uint2 *sp = (const uint2 *)planes;
    return stem_tiles(n, sp, sp + scan_words(n), seed, 0, 1, (uint64_t *)redu);
}

// Longest stem of a changed CDS of an offspring generation (one workgroup),
// seeded by its parent's record: q-gram searches for the runs of at least
// the seed when that is 7 or more (exact), else for the runs of at least a
// guess (exact when one is found), then of at least 7; the tiles follow a
// fruitless guess of 7, a search out of budget, or a CDS the search does not
// take.
__device__ __forceinline__ uint64_t stem_child(int n, const uint2 *planes, int pl, int pp, int pq, int32_t *bcast,
                                               uint64_t *redu, uint16_t *qhead) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint2 *sp = planes, *rp = planes + scan_words(n);
    if (wid == 0) {
        const int seed = stem_seed(sp, n, pl, pp, pq, lane);
        if (lane == 0) {
            bcast[0] = seed;
            bcast[2] = pl;
        }
    }
    __syncthreads();
    const int seed = bcast[0];
    for (int K = qg_first_k(n, seed, bcast[2]); K;) {
        uint64_t qbest;
        if (!stem_qgram(n, sp, rp, K, qhead, bcast, redu, qbest)) break;
        if (seed >= QG_KMIN || (int)(qbest >> 40) >= K) return qbest;
        K = qg_next_k(K);
    }
    return stem_tiles_whole(n, (const lds_uint2 *)planes, seed, (lds_uint64 *)redu);
}

// Workgroups of k_stem per CDS of n bases: about three tiles per lane or
// more, at most STEM_PARTS.
__host__ __device__ inline int stem_parts(int n) {
    const int TT = 2 * tile_prefix(n, tile_blocks(n)), p = TT / (3 * 256);
    return p < 1 ? 1 : (p > STEM_PARTS ? STEM_PARTS : p);
}

// ---------------------------------------------------------------------------
// Evaluation of the CDS flagged dirty, then objectives 0-4
// (calMinimumCAI/CBP/HSC/HD, calMaximumGC3, utils.cuh:271-626), on codon
// codes.  A changed CDS is marked for k_stem (pad = 1); its stem fields keep
// the parent's stem, which seeds the scan (have_parent), or are cleared.  Up
// to four changed CDS share one pass over the positions; the sums are wave
// sums (DPP) added over the waves by one thread per value, which also
// computes the CDS's values (CAI, CPB, HSC quotient, GC3 fraction) kept in
// its record.  Hamming distances take four codons per 32-bit word of two code
// rows, a wave per pair: every pair without a parent, only the pairs of a
// changed CDS with one (sh.hd holds the parent's distances on entry).
// ---------------------------------------------------------------------------
// Per-CDS values of the sums: CAI, CPB, the HSC quotient and the GC3 fraction.
__device__ __forceinline__ void cds_values(CdsStat &s, int L) {
    s.ov[0] = (float)det_exp(((double)s.cai * 0x1p-40) / (double)L);
    s.ov[1] = (float)(((double)s.cpb * 0x1p-40) / (double)(L - 2));
    s.ov[2] = (float)s.hsc / (float)(L - 1);
    s.ov[3] = (float)s.gc3 / (float)L;
}

// The pick-th (0-based) set bit of a candidate mask.
__device__ __forceinline__ int nth_bit(uint32_t mask, uint32_t pick) {
    for (uint32_t q = 0; q < pick; ++q) mask &= mask - 1u;
    return __builtin_ctz(mask);
}

template <class SH>
__device__ void evaluate(const DevJob &J, const DesignLds &D, const uint8_t *cd6, CdsStat *cst, bool have_parent,
                         SH &sh, float *obj, int32_t *oidx) {
    const int L = J.L, m = J.m, n = 3 * L, Ls = D.Ls;
    const int tid = threadIdx.x, NT = blockDim.x, lane = tid & 63, wid = tid >> 6, nw = NT >> 6;
    uint32_t dmask = 0;  // the same in every thread
    for (int c = 0; c < m; ++c) dmask |= (uint32_t)(sh.dirty[c] != 0) << c;
    {
        uint32_t left = dmask;
        while (left) {
            int cs[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                cs[u] = left ? __builtin_ctz(left) : -1;
                left &= left - 1u;
            }
            // one CDS at a time (its sums reduced over the wave into LDS before
            // the next), four codons a thread per round: the pair scores of all
            // four (global table, one dependent load each) are fetched together
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (cs[u] < 0) continue;
                const uint8_t *g = cd6 + cs[u] * Ls;
                int64_t cai = 0, cpb = 0;
                int hsc = 0, gcn = 0;
                for (int k0 = tid; k0 < L; k0 += 4 * NT) {
                    int64_t ps[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int k = k0 + v * NT;
                        const bool pair = k < L - 2;
                        const int kk = pair ? k : 0;  // an index inside the row either way
                        ps[v] = pair ? gload(J.cpfx6 + g[kk] * 64 + g[kk + 1]) : 0;
                    }
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int k = k0 + v * NT;
                        if (k >= L) break;
                        const int a = g[k];
                        cai += D.lw6[a];
                        gcn += code_gc3(a);
                        if (k < L - 1) hsc += code_hidden(D.stop6, a, g[k + 1]);
                        cpb += ps[v];
                    }
                }
                cai = wave_sum_i64(cai);
                cpb = wave_sum_i64(cpb);
                hsc = wave_sum_i32(hsc);
                gcn = wave_sum_i32(gcn);
                if (lane == 0) {
                    sh.red[wid * 8 + 2 * u] = cai;
                    sh.red[wid * 8 + 2 * u + 1] = cpb;
                    sh.red32[wid * 8 + 2 * u] = hsc;
                    sh.red32[wid * 8 + 2 * u + 1] = gcn;
                }
            }
            lds_barrier();
            if (tid < 4) {
                const int c = tid == 0 ? cs[0] : tid == 1 ? cs[1] : tid == 2 ? cs[2] : cs[3];
                if (c >= 0) {
                    int64_t a = 0, b = 0;
                    int h = 0, g = 0;
                    for (int w = 0; w < nw; ++w) {
                        a += sh.red[w * 8 + 2 * tid];
                        b += sh.red[w * 8 + 2 * tid + 1];
                        h += sh.red32[w * 8 + 2 * tid];
                        g += sh.red32[w * 8 + 2 * tid + 1];
                    }
                    CdsStat s = cst[c];
                    if (!have_parent) s.l = s.p = s.q = 0;
                    s.cai = a;
                    s.cpb = b;
                    s.hsc = h;
                    s.gc3 = g;
                    s.pad = 1;
                    s.key = 0;
                    cds_values(s, L);
                    cst[c] = s;
                }
            }
            lds_barrier();
        }
    }
    // pairwise Hamming distances, pairs (i, j), i < j, in lexicographic order,
    // a wave each (pair p on wave p mod nw): every pair without a parent,
    // those with a changed CDS otherwise; the rows' padding codes are 0 in
    // every CDS and add nothing
    {
        const int npair = m * (m - 1) / 2, nwd = Ls >> 2;
        const uint32_t *w6 = (const uint32_t *)cd6;
        int i = 0, j = 1;  // pair wid
        for (int t = 0; t < wid && t < npair; ++t)
            if (++j == m) {
                ++i;
                j = i + 1;
            }
        for (int p = wid; p < npair; p += nw) {
            if (!have_parent || ((dmask >> i) & 1u) || ((dmask >> j) & 1u)) {
                int acc = 0;
                for (int q = lane; q < nwd; q += 64) acc += code_mismatch4(w6[i * nwd + q] ^ w6[j * nwd + q]);
                acc = wave_sum_i32(acc);
                if (lane == 0) sh.hd[p] = acc;
            }
            for (int t = 0; t < nw; ++t)  // the pair nw further on
                if (++j == m) {
                    ++i;
                    j = i + 1;
                }
        }
    }
    lds_barrier();
    // The objectives (calMinimumCAI .. calMaximumGC3, utils.cuh:271-626): each
    // is the largest of the CDS's (pairs') values, ties to the last CDS (pair)
    // -- a serial sweep that takes every value at least as large as the one it
    // holds.  One lane per CDS (pair) and the largest (value bits, index) key
    // over the wave give the same choice (the values are finite).
    if (wid == 0) {
        const float gc3_ref = J.gc3_ref;
        const bool cl = lane < m;
        const int c = cl ? lane : 0;
        auto key = [](bool ok, float v, int i) -> uint64_t {
            return ok ? ((uint64_t)ord_bits(v) << 32) | (uint32_t)i : 0ull;
        };
        const CdsStat &sc = cst[c];
        const uint64_t k_cai = wave_max_u64(key(cl, -sc.ov[0], lane));
        const uint64_t k_cpb = wave_max_u64(key(cl, -sc.ov[1], lane));
        const uint64_t k_hsc = wave_max_u64(key(cl, -sc.ov[2], lane));
        const uint64_t k_gc = wave_max_u64(key(cl, fabsf(sc.ov[3] - gc3_ref), lane));
        const int npair = m * (m - 1) / 2;
        uint64_t kh = 0;
        for (int q = lane; q < npair; q += 64) {
            const uint64_t k = key(true, -((float)sh.hd[q] / (float)n), q);
            kh = k > kh ? k : kh;
        }
        kh = wave_max_u64(kh);
        if (lane < 2 * NOBJ) oidx[lane] = 0;
        if (lane == 0) {
            const int ca = (int)(uint32_t)k_cai, cb = (int)(uint32_t)k_cpb, ch = (int)(uint32_t)k_hsc;
            const int cg = (int)(uint32_t)k_gc, qh = (int)(uint32_t)kh;
            int i = 0, j = 1;  // pair qh
            for (int t = 0; t < qh; ++t)
                if (++j == m) {
                    ++i;
                    j = i + 1;
                }
            oidx[0] = ca;
            oidx[2] = cb;
            oidx[4] = ch;
            oidx[6] = i;
            oidx[7] = j;
            oidx[8] = cg;
            oidx[9] = (cst[cg].ov[3] - gc3_ref) < 0 ? GC_UP : GC_DOWN;
            obj[0] = -cst[ca].ov[0];
            obj[1] = -cst[cb].ov[1];
            obj[2] = -cst[ch].ov[2];
            obj[3] = -((float)sh.hd[qh] / (float)n);
            obj[4] = fabsf(cst[cg].ov[3] - gc3_ref);
        }
    }
    lds_barrier();
}

// ---------------------------------------------------------------------------
// Mutation operators (include/mutation.cuh) on the synonym indices (syn) and
// codon codes (cd6) in LDS.  Each thread owns whole codons; operators that
// read neighbours only touch codons of the other parity or of other CDS, as
// in the reference.  Draws: the i-th slot an operator visits in CDS c takes
// words 2(i&1) (the mutation test) and 2(i&1)+1 (the choice, or the
// shuffle) of Philox(gen, child, c*L + (i >> 1), TAG_MUT), so one call serves
// two slots and a thread takes the two slots of one call.
// ---------------------------------------------------------------------------

struct MutArgs {
    const DevJob *J;
    const DesignLds *D;
    uint8_t *syn, *cd6;
    uint32_t gen, child;
};

__device__ __forceinline__ void set_syn(const MutArgs &A, int c, int k, int s) {
    const int i = c * A.D->Ls + k;
    A.syn[i] = (uint8_t)s;
    A.cd6[i] = A.D->code[A.D->pbase[k] + s];
}

__device__ __forceinline__ void pair_draws(const MutArgs &A, int c, int q, uint32_t r[4]) {
    philox(A.gen, A.child, (uint32_t)(c * A.J->L + q), TAG_MUT, A.J->seed, r);
}

// The codons of an operator whose work per accepted codon is long, in
// batches: wave w tests the draw pairs q = 64w + lane, 64w + 256 + lane, ...
// (test(q) returns bit h set when codon h of pair q passes the operator's
// Bernoulli test), queues the codons that pass (slot 2q + h) in its segment of
// sh.hitq, and runs work(slot) one codon a lane whenever 64 are queued, then on
// the rest.  Accepted codons change only themselves and the tests and the work
// of other codons read only codons no test of this operator accepts, so the
// order of the work does not matter.  Wave-local: the waves never wait for
// each other.
template <class Test, class Work>
__device__ __forceinline__ void hits_by_wave(int npairs, uint16_t *seg, Test test, Work work) {
    const int lane = threadIdx.x & 63;
    int cnt = 0;  // the same in every lane of the wave
    for (int q0 = (int)(threadIdx.x & ~63u); q0 < npairs; q0 += blockDim.x) {
        const int q = q0 + lane;
        const uint32_t hm = q < npairs ? test(q) : 0u;
        const uint64_t b0 = __ballot(hm & 1u), b1 = __ballot(hm & 2u);
        const int r0 = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b0 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b0, 0u));
        const int r1 = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b1 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b1, 0u));
        const int n0 = (int)__popcll(b0);
        if (hm & 1u) seg[cnt + r0] = (uint16_t)(2 * q);
        if (hm & 2u) seg[cnt + n0 + r1] = (uint16_t)(2 * q + 1);
        cnt += n0 + (int)__popcll(b1);
        while (cnt >= 64) {  // at most 63 + 128 queued: within HITQ
            work((int)seg[cnt - 64 + lane]);
            cnt -= 64;
        }
    }
    if (lane < cnt) work((int)seg[lane]);
}

// mutationRandom, mutation.cuh:16-63: every codon of every CDS; a thread
// takes codons 2q and 2q+1 of CDS c for its (c, q) pairs
__device__ void mut_random(const MutArgs &A, int32_t *dirty) {
    const DevJob &J = *A.J;
    const int L = J.L, Ls = A.D->Ls, hL = (L + 1) >> 1, total = J.m * hL;
    int c = (int)((unsigned)threadIdx.x / (unsigned)hL), q = (int)threadIdx.x - c * hL;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        uint32_t r[4];
        pair_draws(A, c, q, r);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = 2 * q + h;
            if (k >= L) break;
            const int ns = A.D->pnum[k];
            if (!bern(r[2 * h], J.thr_pm) || ns == 1) continue;
            const int cur = A.syn[c * Ls + k];
            int nw = (int)uniform_int(r[2 * h + 1], (uint32_t)(ns - 1));
            if (nw >= cur) ++nw;
            set_syn(A, c, k, nw);
            dirty[c] = 1;
        }
        q += blockDim.x;
        while (q >= hL) {
            q -= hL;
            ++c;
        }
    }
}

// mutationCAI with SELECT_UPPER_RANDOM, mutation.cuh:65-142
__device__ void mut_cai(const MutArgs &A, int cds, int32_t *dirty) {
    const DevJob &J = *A.J;
    const int L = J.L, Ls = A.D->Ls;
    for (int q = threadIdx.x; 2 * q < L; q += blockDim.x) {
        uint32_t r[4];
        pair_draws(A, cds, q, r);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = 2 * q + h;
            if (k >= L) break;
            const int ns = A.D->pnum[k];
            if (!bern(r[2 * h], J.thr_pm) || ns == 1) continue;
            const int cur = A.syn[cds * Ls + k];
            if (cur != ns - 1) {
                set_syn(A, cds, k, cur + 1 + (int)uniform_int(r[2 * h + 1], (uint32_t)(ns - 1 - cur)));
                dirty[cds] = 1;
            }
        }
    }
}

// mutationCBP with SELECT_UPPER_RANDOM, mutation.cuh:144-262: the codons
// k = 2i + parity below L-1; a thread takes i = 2q, 2q+1 (k = 4q + parity,
// 4q + parity + 2)
__device__ void mut_cbp(const MutArgs &A, int cds, int32_t *dirty, uint16_t *seg) {
    const DevJob &J = *A.J;
    const int L = J.L, Ls = A.D->Ls, parity = (int)(A.gen & 1u);
    const uint64_t thr = J.thr_2pm;
    const float *const cps6 = J.cps6;
    const uint8_t *row = A.cd6 + cds * Ls;
    const int npairs = L - 1 - parity > 0 ? (L - 1 - parity + 3) >> 2 : 0;  // 4q + parity < L - 1
    auto test = [&](int q) -> uint32_t {
        uint32_t r[4];
        pair_draws(A, cds, q, r);
        uint32_t hm = 0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = 4 * q + parity + 2 * h;
            if (k < L - 1 && bern(r[2 * h], thr) && A.D->pnum[k] > 1) hm |= 1u << h;
        }
        return hm;
    };
    auto work = [&](int slot) {
        const int q = slot >> 1, h = slot & 1, k = 4 * q + parity + 2 * h;
        uint32_t r[4];
        pair_draws(A, cds, q, r);
        const int ns = A.D->pnum[k];
        const int base = A.D->pbase[k], cur = A.syn[cds * Ls + k];
        const int left = (k > 0) ? (int)row[k - 1] : 0;
        const int right = row[k + 1];
        // every synonym's score, the pair-score loads issued together (a
        // missing neighbour adds nothing; stop codes score 0 in the table)
        const bool hl = k > 0, hr = !((A.D->stop6 >> right) & 1u);
        const float *rowl = cps6 + left * 64, *colr = cps6 + right;
        float sc[MAX_SYN];
#pragma unroll
        for (int j = 0; j < MAX_SYN; ++j) {
            const int g = A.D->code[base + (j < ns ? j : 0)];
            const float a = gload(rowl + g), b = gload(colr + g * 64);
            float v = 0.f;
            v += hl ? a : 0.f;
            v += hr ? b : 0.f;
            sc[j] = v;
        }
        float curv = sc[0];
#pragma unroll
        for (int j = 1; j < MAX_SYN; ++j) curv = j == cur ? sc[j] : curv;
        uint32_t cand = 0;  // synonyms scoring above the current codon, in order
#pragma unroll
        for (int j = 0; j < MAX_SYN; ++j)
            if (j < ns && j != cur && sc[j] > curv) cand |= 1u << j;
        if (cand) {
            set_syn(A, cds, k, nth_bit(cand, uniform_int(r[2 * h + 1], (uint32_t)__popc(cand))));
            dirty[cds] = 1;
        }
    };
    hits_by_wave(npairs, seg, test, work);
}

// mutationHSC with SELECT_UPPER_RANDOM, mutation.cuh:264-388: the codons
// k = 2i + parity below L, two per thread as in mut_cbp
__device__ void mut_hsc(const MutArgs &A, int cds, int32_t *dirty, uint16_t *seg) {
    const DevJob &J = *A.J;
    const int L = J.L, Ls = A.D->Ls, parity = (int)(A.gen & 1u);
    const uint64_t thr = J.thr_2pm;
    const uint8_t *row = A.cd6 + cds * Ls;
    const uint64_t stop6 = A.D->stop6;
    const int npairs = L - parity > 0 ? (L - parity + 3) >> 2 : 0;  // 4q + parity < L
    auto test = [&](int q) -> uint32_t {
        uint32_t r[4];
        pair_draws(A, cds, q, r);
        uint32_t hm = 0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = 4 * q + parity + 2 * h;
            if (k < L && bern(r[2 * h], thr) && A.D->pnum[k] > 1) hm |= 1u << h;
        }
        return hm;
    };
    auto work = [&](int slot) {
        const int q = slot >> 1, h = slot & 1, k = 4 * q + parity + 2 * h;
        uint32_t r[4];
        pair_draws(A, cds, q, r);
        const int ns = A.D->pnum[k];
        const int base = A.D->pbase[k], cur = A.syn[cds * Ls + k];
        const int left = (k > 0) ? (int)row[k - 1] : -1;
        const int right = (k < L - 1) ? (int)row[k + 1] : -1;
        auto count = [&](int j) -> int {
            const int g = A.D->code[base + j];
            int v = 0;
            if (left >= 0) v += code_hidden(stop6, left, g);
            if (right >= 0) v += code_hidden(stop6, g, right);
            return v;
        };
        const int curv = count(cur);
        uint32_t cand = 0;
        for (int j = 0; j < ns; ++j)
            if (j != cur && count(j) > curv) cand |= 1u << j;
        if (cand) {
            set_syn(A, cds, k, nth_bit(cand, uniform_int(r[2 * h + 1], (uint32_t)__popc(cand))));
            dirty[cds] = 1;
        }
    };
    hits_by_wave(npairs, seg, test, work);
}

// indexArrayShuffling, utils.cuh:203-218: Fisher-Yates over the n <= 6
// synonyms, entries i = n-1 .. 1 swapped with j_i, the mixed-radix digits
// (radix i+1) of uniform_int(x, n!).  Digits by multiplication: for r < 720,
// (r * (65536/d + 1)) >> 16 is r / d for d = 2..6.  The order is packed in
// 4-bit fields (entry i at bits 4i) so that it stays in registers.
__device__ __forceinline__ uint32_t shuffle(uint32_t x, int n) {
    const uint32_t fact = n == 6 ? 720u : n == 5 ? 120u : n == 4 ? 24u : n == 3 ? 6u : n == 2 ? 2u : 1u;
    uint32_t r = uniform_int(x, fact);
    uint32_t perm = 0x543210u;
#pragma unroll
    for (int i = 5; i > 0; --i) {
        if (i >= n) continue;
        const uint32_t d = (uint32_t)(i + 1), q = (r * (65536u / d + 1u)) >> 16, j = r - q * d;
        r = q;
        const uint32_t vi = (perm >> (4 * i)) & 0xFu, vj = (perm >> (4 * j)) & 0xFu;
        perm &= ~((0xFu << (4 * i)) | (0xFu << (4 * j)));
        perm |= (vj << (4 * i)) | (vi << (4 * j));
    }
    return perm;
}
__device__ __forceinline__ int perm_at(uint32_t perm, int j) { return (int)((perm >> (4 * j)) & 0xFu); }

// mutationHD, mutation.cuh:390-534: codons of cds1, scored against all CDS
__device__ void mut_hd(const MutArgs &A, int cds1, int cds2, int32_t *dirty, uint16_t *seg) {
    const DevJob &J = *A.J;
    const int L = J.L, m = J.m, Ls = A.D->Ls;
    const uint64_t thr = J.thr_pm;
    auto test = [&](int q) -> uint32_t {
        uint32_t r[4];
        pair_draws(A, cds1, q, r);
        uint32_t hm = 0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = 2 * q + h;
            if (k < L && bern(r[2 * h], thr) && A.D->pnum[k] > 1) hm |= 1u << h;
        }
        return hm;
    };
    auto work = [&](int k) {
        uint32_t r[4];
        pair_draws(A, cds1, k >> 1, r);
        const int ns = A.D->pnum[k];
        const uint32_t order = shuffle(r[2 * (k & 1) + 1], ns);
        const int base = A.D->pbase[k], cur = A.syn[cds1 * Ls + k];
        const int g2 = A.cd6[cds2 * Ls + k];
        // every candidate's codes (the synonyms in shuffled order, slot
        // MAX_SYN the current codon) first, then the smallest distance
        // to the other CDS's codons at k (only cds1's codons change here,
        // each by its own lane) and the distance to cds2's codon
        int gc[MAX_SYN + 1], mn[MAX_SYN + 1], pr[MAX_SYN + 1];
#pragma unroll
        for (int j = 0; j <= MAX_SYN; ++j) {
            gc[j] = A.D->code[base + (j == MAX_SYN ? cur : perm_at(order, j < ns ? j : ns - 1))];
            mn[j] = 127;
            pr[j] = code_mismatch(gc[j], g2);
        }
        for (int o = 0; o < m; ++o) {
            if (o == cds1) continue;
            const int oc = A.cd6[o * Ls + k];
#pragma unroll
            for (int j = 0; j <= MAX_SYN; ++j) mn[j] = min(mn[j], code_mismatch(gc[j], oc));
        }
        const int cur_pair = pr[MAX_SYN], cur_mhd = mn[MAX_SYN];
        int best_pair = -1, best_mhd = -1, best = -1;
#pragma unroll
        for (int j = 0; j < MAX_SYN; ++j) {
            const int sh = perm_at(order, j);
            if (j >= ns || sh == cur) continue;
            const int np_ = pr[j], nm = mn[j];
            if (nm > cur_mhd && nm > best_mhd) {
                best_mhd = nm;
                best = sh;
            } else if (best_mhd == -1 && nm == cur_mhd && np_ > cur_pair && np_ > best_pair) {
                best_pair = np_;
                best = sh;
            }
        }
        if (best_mhd != -1 || best_pair != -1) {
            set_syn(A, cds1, k, best);
            dirty[cds1] = 1;
        }
    };
    hits_by_wave((L + 1) >> 1, seg, test, work);
}

// mutationGC3, mutation.cuh:536-653 (probability = the parent's GC3 deviation)
__device__ void mut_gc3(const MutArgs &A, int cds, int dir, float dev, int32_t *dirty, uint16_t *seg) {
    const DevJob &J = *A.J;
    const int L = J.L, Ls = A.D->Ls;
    const uint64_t thr = bern_threshold(dev);
    auto test = [&](int q) -> uint32_t {
        uint32_t r[4];
        pair_draws(A, cds, q, r);
        uint32_t hm = 0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = 2 * q + h;
            if (k < L && bern(r[2 * h], thr) && A.D->pnum[k] > 1) hm |= 1u << h;
        }
        return hm;
    };
    auto work = [&](int k) {
        uint32_t r[4];
        pair_draws(A, cds, k >> 1, r);
        const int ns = A.D->pnum[k];
        const uint32_t order = shuffle(r[2 * (k & 1) + 1], ns);
        const int base = A.D->pbase[k], cur = A.syn[cds * Ls + k];
        const int curg = code_gc3(A.D->code[base + cur]);
        for (int j = 0; j < ns; ++j) {
            const int sh = perm_at(order, j);
            if (sh == cur) continue;
            const int ng = code_gc3(A.D->code[base + sh]);
            if ((dir == GC_UP && ng > curg) || (dir == GC_DOWN && ng < curg)) {
                set_syn(A, cds, k, sh);
                dirty[cds] = 1;
                break;
            }
        }
    };
    hits_by_wave((L + 1) >> 1, seg, test, work);
}

// mutationSL, mutation.cuh:775-969.  The accept test of the reference
// (mutation.cuh:843,896,947) compares the parent's SL objective with the
// integer quotient l_new / ((n-4)/2); it only needs to know whether the CDS
// still has any stem and whether it is one full-length hairpin, which two
// O(n) workgroup checks over the CDS's codes decide.
template <class SH>
__device__ void mut_sl(const MutArgs &A, SH &sh, int32_t *dirty) {
    const DevJob &J = *A.J;
    const int L = J.L, n = 3 * L, Ls = A.D->Ls, tid = threadIdx.x, NT = blockDim.x;
    const int cds = sh.pidx[10];
    const uint8_t *row = A.cd6 + cds * Ls;
    const int P0 = sh.ppql[0], Lst = sh.ppql[2];
    const int l_idx = P0 / 3, r_idx = (P0 + Lst - 1) / 3;
    const int mid = (r_idx + l_idx) / 2;
    const int stem_len = r_idx - l_idx + 1, half = stem_len / 2;
    const bool odd = (stem_len % 2) == 1;
    const int l_st0 = odd ? mid - 1 : mid, r_st0 = mid + 1;
    const int den = J.den_sl;
    const float parent_sl = sh.pobj[5];
    const int trials = (odd ? 1 : 0) + 2 * half;
    const int lane = tid & 63, wid = tid >> 6, nw = NT >> 6;
    int *ctl = sh.misc + 11;  // [0] changed, [1] saved, [2] terminate, [3] cur_L
    if (tid == 0) ctl[3] = Lst;
    lds_barrier();
    for (int t = 0; t < trials; ++t) {
        int pos;
        if (odd && t == 0) {
            pos = mid;
        } else {
            const int u = t - (odd ? 1 : 0);
            pos = (u & 1) ? r_st0 + (u >> 1) : l_st0 - (u >> 1);
        }
        if (tid == 0) {
            uint32_t r[4];
            philox(A.gen, A.child, (uint32_t)t, TAG_SL, J.seed, r);
            const int ns = A.D->pnum[pos];
            ctl[0] = 0;
            ctl[1] = A.syn[cds * Ls + pos];
            if (bern(r[0], J.thr_pm) && ns > 1) {
                const int cur = A.syn[cds * Ls + pos];
                int nw = (int)uniform_int(r[1], (uint32_t)(ns - 1));
                if (nw == cur) ++nw;  // mutation.cuh:827 compares with ==
                set_syn(A, cds, pos, nw);
                ctl[0] = 1;
            }
        }
        lds_barrier();
        // first and last position of each base seen by this thread (positions
        // rise, so the first hit is the first); kept in registers, one per base
        int f[4] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF}, l4[4] = {-1, -1, -1, -1};
        int okA = 1, okB = 1, okC = 1;
        for (int i = tid; i < n; i += NT) {
            const int b = code_base(row, i);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (b == q) {
                    f[q] = min(f[q], i);
                    l4[q] = i;
                }
        }
        for (int i = tid; i < den; i += NT) {
            okA &= ((code_base(row, i) ^ code_base(row, n - 1 - i)) == 3);
            if (n & 1) {
                okB &= ((code_base(row, i) ^ code_base(row, n - 2 - i)) == 3);
                okC &= ((code_base(row, 1 + i) ^ code_base(row, n - 1 - i)) == 3);
            }
        }
        // wave minima, maxima and conjunctions (DPP, ballots), combined over
        // the waves by thread 0
        const bool allA = __ballot(!okA) == 0, allB = __ballot(!okB) == 0, allC = __ballot(!okC) == 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            f[q] = wave_min_i32(f[q]);
            l4[q] = wave_max_i32(l4[q]);
        }
        if (lane == 0) {
            int *w = sh.slw + wid * 11;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                w[q] = f[q];
                w[4 + q] = l4[q];
            }
            w[8] = allA;
            w[9] = allB;
            w[10] = allC;
        }
        lds_barrier();
        if (tid == 0) {
            int first[4] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF}, last[4] = {-1, -1, -1, -1};
            int full[3] = {1, 1, 1};
            for (int w = 0; w < nw; ++w) {
                const int *v = sh.slw + w * 11;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    first[q] = min(first[q], v[q]);
                    last[q] = max(last[q], v[4 + q]);
                }
#pragma unroll
                for (int q = 0; q < 3; ++q) full[q] &= v[8 + q];
            }
            bool any = false;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int b = a ^ 3;
                if (first[a] != 0x7FFFFFFF && last[b] >= 0 && last[b] - first[a] > MIN_HAIRPIN) any = true;
            }
            bool is_full = false;
            if (den > 0) {
                if (full[0] && (n - 1 - 2 * (den - 1)) > MIN_HAIRPIN) is_full = true;
                if ((n & 1) && (full[1] || full[2]) && (n - 2 - 2 * (den - 1)) > MIN_HAIRPIN) is_full = true;
            }
            if (any) ctl[3] = is_full ? den : 0;
            const float rhs = (float)(den > 0 ? ctl[3] / den : 0);
            ctl[2] = parent_sl > rhs;
            if (!ctl[2] && ctl[0]) set_syn(A, cds, pos, ctl[1]);
            if (ctl[2] && ctl[0]) dirty[cds] = 1;
        }
        lds_barrier();
        if (ctl[2]) break;
    }
}

// Bit planes of the changed CDS into the job's plane buffer: bit p of the
// sequence plane is position p, of the reverse-complement plane position
// n-1-p complemented; positions past n are zero; a word is (low bits, high
// bits) of 32 positions.  The forward words are built 32 codons (three words)
// per lane from the codes, four codons per 32-bit step, and staged in LDS
// (stage + c * stride for CDS c); the reverse-complement word j is the
// forward window of bits n-32-32j .. n-1-32j, bit-reversed and inverted.
// rows: the CDS's code rows (stride Ls); dl: the changed CDS, nd of them.
__device__ void build_planes(const uint8_t *rows, int Ls, int L, const int *dl, int nd, uint2 *stage, int stride,
                             uint2 *gbase) {
    const int n = 3 * L, W = scan_words(n), ng = (W + 2) / 3, tid = threadIdx.x, NT = blockDim.x;
    const int nwd = Ls >> 2;
    for (int i = tid; i < nd * ng; i += NT) {
        const int e = i / ng, g = i - e * ng, c = dl[e];
        const uint32_t *row = (const uint32_t *)(rows + c * Ls);
        uint32_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
        for (int t = 0; t < 8; ++t) {  // codons 32g + 4t .. 32g + 4t + 3 (zero past the row)
            const int q = 8 * g + t;
            const uint32_t d = q < nwd ? row[q] : 0u;
            // per byte the three bases' low (high) bits in position order
            const uint32_t l = ((d >> 4) & 0x01010101u) | ((d >> 1) & 0x02020202u) | ((d << 2) & 0x04040404u);
            const uint32_t h = ((d >> 5) & 0x01010101u) | ((d >> 2) & 0x02020202u) | ((d << 1) & 0x04040404u);
            const uint32_t l12 = (l & 7u) | ((l >> 5) & 0x38u) | ((l >> 10) & 0x1C0u) | ((l >> 15) & 0xE00u);
            const uint32_t h12 = (h & 7u) | ((h >> 5) & 0x38u) | ((h >> 10) & 0x1C0u) | ((h >> 15) & 0xE00u);
            const int b = 12 * t, w = b >> 5, o = b & 31;  // compile-time after unrolling
            lo[w] |= l12 << o;
            hi[w] |= h12 << o;
            if (o > 20) {
                lo[w + 1] |= l12 >> (32 - o);
                hi[w + 1] |= h12 >> (32 - o);
            }
        }
        uint2 *st = stage + (size_t)e * stride;
#pragma unroll
        for (int u = 0; u < 3; ++u)
            if (3 * g + u < W) st[3 * g + u] = make_uint2(lo[u], hi[u]);
        if (g == 0) st[W] = make_uint2(0u, 0u);
    }
    lds_barrier();
    for (int i = tid; i < nd * W; i += NT) {
        const int e = i / W, j = i - e * W, c = dl[e];
        const uint2 *st = stage + (size_t)e * stride;
        uint2 *sp = gbase + (size_t)c * 2 * W, *rp = sp + W;
        gstore((uint64_t *)(sp + j), (uint64_t)st[j].x | ((uint64_t)st[j].y << 32));
        const int s0 = n - 32 - 32 * j;  // first bit of the forward window
        uint32_t wl = 0, wh = 0;
        if (s0 >= 0) {
            const uint2 a = st[s0 >> 5], b = st[(s0 >> 5) + 1];
            wl = __builtin_amdgcn_alignbit(b.x, a.x, (uint32_t)(s0 & 31));
            wh = __builtin_amdgcn_alignbit(b.y, a.y, (uint32_t)(s0 & 31));
        } else if (s0 > -32) {
            wl = st[0].x << (-s0);
            wh = st[0].y << (-s0);
        }
        const int cnt = n - 32 * j;  // positions of word j below n
        const uint32_t mask = cnt >= 32 ? ~0u : (cnt <= 0 ? 0u : (1u << cnt) - 1u);
        const uint32_t rl = ~__builtin_bitreverse32(wl) & mask, rh = ~__builtin_bitreverse32(wh) & mask;
        gstore((uint64_t *)(rp + j), (uint64_t)rl | ((uint64_t)rh << 32));
    }
}

// The index of the c-th survivor (0-based, ascending pool index) in the
// survivor mask of W words: wave-uniform, every lane of the wave takes part.
__device__ int survivor_at(const uint64_t *selm, int W, int c, uint64_t first) {
    const int lane = threadIdx.x & 63;
    int base = 0;
    uint64_t x = lane < W ? first : 0ull;  // (word `lane` of the mask, loaded by the caller)
    for (int w0 = 0; w0 < W; w0 += 64) {
        if (w0) {
            const int w = w0 + lane;
            x = w < W ? gload(selm + w) : 0ull;
        }
        const int incl = wave_scan_i32((int)__popcll(x));  // inclusive prefix of the words' counts
        const uint64_t hit = __ballot(base + incl > c);
        if (hit) {
            const int q = (int)__builtin_ctzll(hit);  // wave-uniform: read lane q directly
            uint64_t y = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(x >> 32), q) << 32) |
                         (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, q);
            int k = c - (base + __builtin_amdgcn_readlane(incl, q) - (int)__popcll(y));  // rank within the word
            int pos = 0;
#pragma unroll
            for (int h = 32; h > 0; h >>= 1) {  // the k-th set bit of y
                const uint64_t low = (1ull << h) - 1ull;
                const int n = (int)__popcll(y & low);
                if (k >= n) {
                    k -= n;
                    y >>= h;
                    pos += h;
                } else {
                    y &= low;
                }
            }
            return (w0 + q) * 64 + pos;
        }
        base += __builtin_amdgcn_readlane(incl, 63);
    }
    return 0;  // fewer than c + 1 survivors: not reached
}

// ---------------------------------------------------------------------------
// k_design: one workgroup per design.
//   INIT       random genome (design 0: best-CAI codons), genPopulation,
//              utils.cuh:1047-1104, then evaluation into buffer 0
//   EVAL       evaluation of genome[0][d] into buffer 0
//   OFFSPRING  survivor copy, mutation, evaluation (mutationKernel)
// A design's genome row is m CDS rows of stride Ls; it moves between HBM and
// LDS in 16-byte words.
// ---------------------------------------------------------------------------
template <int MODE, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6))) void k_design(const DevJob *__restrict__ jobs, const int2 *__restrict__ items,
                                                uint32_t *__restrict__ stem_list, int32_t *stem_ctl, int stem_group,
                                                int stem_cap, int item_base) {
    extern __shared__ __align__(16) uint8_t dyn[];
    __shared__ DesignShared<NT / 64> sh;
    const int2 it = items[blockIdx.x];
    const DevJob &J = jobs[it.x];
    const int L = J.L, m = J.m, Ls = J.Ls, N = J.N;
    const int tid = threadIdx.x;
    const int G16 = (m * Ls) >> 4;  // 16-byte words of a design row
    const int SA = syn_stride(L);
    uint8_t *pbase = dyn + DL_POS;
    uint8_t *pnum = pbase + Ls;
    uint8_t *syn = pnum + Ls;       // rows at stride Ls; the planes' stage afterwards (stride SA)
    uint8_t *cd6 = syn + m * SA;
    CdsStat *cst = (CdsStat *)(cd6 + m * Ls);
    const int npair = m * (m - 1) / 2;
    const DesignLds D{(const int64_t *)(dyn + DL_LW6), dyn + DL_CODE, pbase, pnum, J.stop6, Ls};

    // The descriptor fields of the staging phase, read before its stores (a
    // store through a pool pointer could alias the descriptor as far as the
    // compiler knows, and would make every later field a reload).
    const uint8_t *const pos = J.pos;
    const int64_t *const lw6 = J.lw6;
    const uint8_t *const ctab = J.code;
    int8_t *const genome0 = J.genome[0], *const genome1 = J.genome[1];
    float *const obj0 = J.obj[0], *const obj1 = J.obj[1];
    int8_t *const oidx0 = J.oidx[0], *const oidx1 = J.oidx[1];
    int32_t *const pql0 = J.pql[0], *const pql1 = J.pql[1];
    CdsStat *const cst0 = J.cst[0], *const cst1 = J.cst[1];
    int32_t *const hdp0 = J.hdp[0], *const hdp1 = J.hdp[1];
    int32_t *const np = J.np;
    SelAcc *const acc = J.acc;
    // stage the tables and the synonym ranges (rows pbase and pnum, as the
    // job keeps them: padding positions have one synonym at codon 0)
    // Every load of the set-up is issued before the first LDS store (a store
    // under a branch would wait for its own load first): the tables, the
    // first words of the synonym ranges, the generation counter and the first
    // words of the survivor mask.
    const int npw = (2 * Ls) >> 4;  // 16-byte words of the rows pbase and pnum
    const uint8_t *p_pos = pos, *p_code = ctab;
    const int64_t *p_lw = lw6;
    const uint64_t *p_sel = J.selm;
    const SelState *p_st = J.st;
    int W = J.W;
    asm volatile("" : "+s"(p_pos), "+s"(p_code), "+s"(p_lw), "+s"(p_sel), "+s"(p_st), "+s"(W));
    int64_t v_lw = gload(p_lw + (tid & 63));
    uint32_t v_code = gload(p_code + (tid & 63));
    u32x4 v_pos = gload((const u32x4 *)p_pos + min(tid, npw - 1));
    uint32_t g0 = 0;
    uint64_t sel0 = 0;
    if (MODE == 2) {
        g0 = gload(&p_st->gen);
        sel0 = gload(p_sel + min(tid & 63, W - 1));
    }
    // (taken here, so that no load is moved into the branches below)
    asm volatile("" : "+v"(v_lw), "+v"(v_code), "+v"(v_pos), "+v"(sel0));
    if (tid < 64)
        ((int64_t *)(dyn + DL_LW6))[tid] = v_lw;
    else if (tid < 128)
        dyn[DL_CODE + tid - 64] = (uint8_t)v_code;
    if (tid < npw) ((u32x4 *)pbase)[tid] = v_pos;
    for (int i = tid + NT; i < npw; i += NT) ((u32x4 *)pbase)[i] = gload((const u32x4 *)pos + i);
    if (tid < MAX_M) sh.dirty[tid] = (MODE != 2);
    uint32_t gen = 0;
    int dst = 0, out_slot = it.y, src = 0, parent = 0;
    if (MODE == 2) {
        gen = __builtin_amdgcn_readfirstlane(g0);
        src = (int)(gen & 1u);
        dst = src ^ 1;
        out_slot = N + it.y;
        parent = survivor_at(p_sel, W, it.y, sel0);
    }
    u32x4 *syn4 = (u32x4 *)syn;
    if (MODE == 0) {
        // positions take their synonym counts from the job's rows (the LDS
        // copy above is not yet visible to every thread)
        const int d = it.y;
        for (int c = 0; c < m; ++c)
            for (int k = tid; k < Ls; k += NT) {
                int s = 0;
                if (k >= L) {
                    s = 0;
                } else if (d == 0) {
                    s = (int)gload(pos + Ls + k) - 1;
                } else {
                    uint32_t r[4];
                    philox(0u, (uint32_t)d, (uint32_t)(c * L + k), TAG_INIT, J.seed, r);
                    s = (int)uniform_int(r[0], (uint32_t)gload(pos + Ls + k));
                }
                syn[c * Ls + k] = (uint8_t)s;
            }
    } else if (MODE == 1) {
        const u32x4 *sg = (const u32x4 *)(genome0 + (size_t)it.y * m * Ls);
        for (int i = tid; i < G16; i += NT) syn4[i] = gload(sg + i);
    } else {
        const int c = it.y;
        // the survivor copy (copySolution, utils.cuh:1106-1147): every load
        // is issued before the stores, at clamped indices
        const float *const obj_s = src ? obj1 : obj0;
        float *const obj_d = src ? obj0 : obj1;
        const int8_t *const oidx_s = src ? oidx1 : oidx0;
        int8_t *const oidx_d = src ? oidx0 : oidx1;
        const int32_t *const pql_s = src ? pql1 : pql0;
        int32_t *const pql_d = src ? pql0 : pql1;
        const CdsStat *const cst_s = src ? cst1 : cst0;
        CdsStat *const cst_d = src ? cst0 : cst1;
        const int32_t *const hdp_s = src ? hdp1 : hdp0;
        int32_t *const hdp_d = src ? hdp0 : hdp1;
        const u32x4 *pg = (const u32x4 *)((src ? genome1 : genome0) + (size_t)parent * m * Ls);
        u32x4 *sg = (u32x4 *)((dst ? genome1 : genome0) + (size_t)c * m * Ls);
        // the records' loads first (lanes of wave 0), then the genome row's
        float vo = 0.f;
        int8_t vi = 0;
        int32_t vp = 0;
        CdsStat cs{};
        int32_t vh = 0;
        const int hp = tid - 64;  // lanes 64 .. 64 + npair - 1: the pairs' distances
        if (hp >= 0 && hp < npair) vh = gload(hdp_s + (size_t)parent * npair + hp);
        if (tid < 16) {
            vo = gload(obj_s + (size_t)parent * NOBJ + min(tid, NOBJ - 1));
            vi = gload(oidx_s + (size_t)parent * 2 * NOBJ + min(tid, 2 * NOBJ - 1));
            vp = gload(pql_s + (size_t)parent * 4 + min(tid, 3));
            cs = gload_stat(cst_s + (size_t)parent * m + min(tid, m - 1));
        }
        for (int i0 = tid; i0 < G16; i0 += 4 * NT) {
            u32x4 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = gload(pg + min(i0 + u * NT, G16 - 1));
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (i0 + u * NT < G16) {
                    gstore(sg + i0 + u * NT, v[u]);
                    syn4[i0 + u * NT] = v[u];
                }
        }
        if (tid < 16) {
            if (tid < NOBJ) {
                gstore(obj_d + (size_t)c * NOBJ + tid, vo);
                sh.pobj[tid] = vo;
            }
            if (tid < 2 * NOBJ) {
                gstore(oidx_d + (size_t)c * 2 * NOBJ + tid, vi);
                sh.pidx[tid] = vi;
            }
            if (tid < 4) {
                gstore(pql_d + (size_t)c * 4 + tid, vp);
                sh.ppql[tid] = vp;
            }
            if (tid < m) {
                cs.pad = 0;
                gstore_stat(cst_d + (size_t)c * m + tid, cs);
                cst[tid] = cs;
            }
        }
        if (hp >= 0 && hp < npair) {
            gstore(hdp_d + (size_t)c * npair + hp, vh);
            sh.hd[hp] = vh;
        }
        if (tid == NT - 1) {  // the sort of this generation starts from cleared counts
            gstore(np + c, 0);
            gstore(np + N + c, 0);
        }
        if (c == 0 && tid >= 128 && tid < 128 + ACC_WORDS) gstore((uint32_t *)acc + tid - 128, acc_clear_word(tid - 128));
    }
    lds_barrier();  // tables, synonym ranges, genome row and records
    // codon codes of the row, four per thread step; the padding positions
    // are forced to synonym 0 and code 0 (the table reads stay below 64: a
    // padding position reads codon 0)
    {
        const int nwd = Ls >> 2;
        const uint32_t *pb4 = (const uint32_t *)pbase;
        uint32_t *s4 = (uint32_t *)syn, *c4 = (uint32_t *)cd6;
        for (int c = 0; c < m; ++c)
            for (int q = tid; q < nwd; q += NT) {
                const int left = L - 4 * q;  // positions of the word below L
                const uint32_t keep = left >= 4 ? ~0u : (left <= 0 ? 0u : (1u << (8 * left)) - 1u);
                const uint32_t s = s4[c * nwd + q] & keep;
                const uint32_t g = s + pb4[q];  // bytewise: synonyms < 6, first codons < 64
                uint32_t code = 0;
#pragma unroll
                for (int b = 0; b < 4; ++b) code |= (uint32_t)D.code[(g >> (8 * b)) & 0xFFu] << (8 * b);
                s4[c * nwd + q] = s;
                c4[c * nwd + q] = code & keep;
            }
    }
    lds_barrier();
    if (MODE == 2) {
        const int c = it.y;
        uint32_t r[4];
        philox(gen, (uint32_t)c, 0u, TAG_TYPE, J.seed, r);
        const int type = (int)uniform_int(r[0], (uint32_t)(NOBJ + 1));  // main.cu:172-178
        const MutArgs A{&J, &D, syn, cd6, gen, (uint32_t)c};
        switch (type) {  // main.cu:181-211
        case 0: mut_random(A, sh.dirty); break;
        case 1: mut_cai(A, sh.pidx[0], sh.dirty); break;
        case 2: mut_cbp(A, sh.pidx[2], sh.dirty, sh.hitq[tid >> 6]); break;
        case 3: mut_hsc(A, sh.pidx[4], sh.dirty, sh.hitq[tid >> 6]); break;
        case 4: mut_hd(A, sh.pidx[6], sh.pidx[7], sh.dirty, sh.hitq[tid >> 6]); break;
        case 5:
            if (sh.pidx[9] == GC_UP || sh.pidx[9] == GC_DOWN)
                mut_gc3(A, sh.pidx[8], sh.pidx[9], sh.pobj[4], sh.dirty, sh.hitq[tid >> 6]);
            break;
        default: mut_sl(A, sh, sh.dirty); break;
        }
        lds_barrier();
    }

    float *oobj = (dst ? obj1 : obj0) + (size_t)out_slot * NOBJ;
    __shared__ float s_obj[NOBJ];
    __shared__ int32_t s_idx[2 * NOBJ];
    // the design's genome row is final: its stores drain during the evaluation
    u32x4 *og = (u32x4 *)((dst ? genome1 : genome0) + (size_t)out_slot * m * Ls);
    for (int i = tid; i < G16; i += NT) gstore(og + i, syn4[i]);
    evaluate(J, D, cd6, cst, MODE == 2, sh, s_obj, s_idx);
    {  // planes of the changed CDS for k_stem, staged where the syn rows were
        int nd = 0;
        for (int c = 0; c < m; ++c) {
            if (tid == 0 && sh.dirty[c]) sh.dlist[nd] = c;
            nd += sh.dirty[c] != 0;
        }
        lds_barrier();  // the syn rows are stored and the list is written
        build_planes(cd6, Ls, L, sh.dlist, nd, (uint2 *)syn, SA / 8,
                     J.planes + (size_t)out_slot * m * 2 * scan_words(3 * L));
    }
    if (tid < NOBJ - 1) gstore(oobj + tid, s_obj[tid]);  // objective 5: k_stem_finish
    if (tid < 2 * NOBJ && tid != 10) gstore((dst ? oidx1 : oidx0) + (size_t)out_slot * 2 * NOBJ + tid, (int8_t)s_idx[tid]);
    if (tid < m) gstore_stat((dst ? cst1 : cst0) + (size_t)out_slot * m + tid, cst[tid]);
    if (tid < npair) gstore((dst ? hdp1 : hdp0) + (size_t)out_slot * npair + tid, sh.hd[tid]);
    // CDS evaluated (k_stem scans the same ones); the unchanged CDS's longest
    // stem.  The first lane of the last wave does this: a wave's wait for the
    // list atomic's result also waits for the wave's earlier stores, and the
    // rows, planes and records go to the first waves first.
    if (tid == NT - 64) {
        unsigned long long nd = 0;
        int lmax = 0;
        for (int c = 0; c < m; ++c) {
            nd += sh.dirty[c] != 0;
            if (!sh.dirty[c]) lmax = max(lmax, cst[c].l);
        }
        gstore(J.stem_max + out_slot, lmax);  // (k_dominance counts the offspring's changed CDS)
        if (MODE == MODE_OFFSPRING && nd) {  // the changed CDS for k_stem_list, in the child's sublist
            const int item = item_base + (int)blockIdx.x, k = item / stem_group;
            int e = atomicAdd(stem_ctl + k, (int)nd);
            for (int c = 0; c < m; ++c)
                if (sh.dirty[c] && e < stem_cap)
                    gstore(stem_list + (size_t)k * stem_cap + e++, ((uint32_t)item << 4) | (uint32_t)c);
        }
    }
}

// Buffer and slot of the design an evaluation work item refers to: children
// of the current generation sit at N + child in the next buffer.
__device__ __forceinline__ void item_slot(const DevJob &J, int mode, int y, int &buf, int &slot) {
    if (mode == MODE_OFFSPRING) {
        buf = (int)((J.st->gen + 1u) & 1u);
        slot = J.N + y;
    } else {
        buf = 0;
        slot = y;
    }
}

// ---------------------------------------------------------------------------
// k_stem (initialisation and evaluation): up to STEM_PARTS workgroups per
// (work item, CDS); blockIdx.x = part * nic + item * mg + c with mg the
// largest m of the jobs launched and nic = items * mg.  Part-major order keeps
// the workgroups that do work consecutive, so the round-robin placement of
// workgroups on the XCDs spreads them over all eight (a CDS-major order with
// unused parts would leave the working ones on the XCDs of one residue
// class).  CDS not marked by k_design (pad 0) are skipped.  Each part folds
// its best key into the record with an atomic maximum; k_stem_finish decodes
// it.
// ---------------------------------------------------------------------------
// Six waves per SIMD (at most 80 VGPRs) keep ~1500 workgroups resident at
// once on 256 CUs.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_stem(const DevJob *__restrict__ jobs, const int2 *__restrict__ items, int mode,
                                              int mg, int nic) {
    extern __shared__ __align__(16) uint8_t dyn[];
    __shared__ uint64_t redu[16];
    const int part = (int)(blockIdx.x / (unsigned)nic), ic = (int)(blockIdx.x % (unsigned)nic);
    const int2 it = items[ic / mg];
    const int c = ic % mg;
    const DevJob &J = jobs[it.x];
    const int L = J.L, m = J.m, n = 3 * L, parts = stem_parts(n);
    if (c >= m || part >= parts) return;
    int buf, slot;
    item_slot(J, mode, it.y, buf, slot);
    CdsStat *cs = J.cst[buf] + (size_t)slot * m + c;
    if (cs->pad == 0) return;  // same word for every thread; k_stem never writes it
    uint64_t *planes = (uint64_t *)dyn;  // the CDS's planes as k_design left them
    const int nwords = 2 * scan_words(n);
    const uint64_t *gp = (const uint64_t *)(J.planes + ((size_t)slot * m + c) * nwords);
    for (int i = threadIdx.x; i < nwords; i += blockDim.x) planes[i] = gload(gp + i);
    __syncthreads();
    const uint2 *sp = (const uint2 *)planes;
    const uint64_t key = stem_tiles(n, sp, sp + scan_words(n), 0, part, parts, redu);
    if (threadIdx.x == 0 && key) atomicMax((unsigned long long *)&cs->key, (unsigned long long)key);
}

// ---------------------------------------------------------------------------
// k_stem_list (offspring generations): the changed CDS k_design listed,
// entry (item << 4) | c, in `groups` sublists of cap entries with their
// counts in ctl; the entries are numbered sublist by sublist (a prefix sum
// of the counts), which keeps the children list's order, and gridDim.x
// workgroups take them round-robin, one workgroup per CDS.  The parent's stem seeds the
// q-gram search (stem_child).  Each entry's key goes into the CDS record and
// its length into the design's longest stem (stem_max) with atomic maxima.
// k_dominance clears ctl[0] for the next generation.
// ---------------------------------------------------------------------------
template <int WAVES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_stem_list(
    const DevJob *__restrict__ jobs, const int2 *__restrict__ items, const uint32_t *__restrict__ list,
    const int32_t *__restrict__ ctl, int groups, int cap) {
    extern __shared__ __align__(16) uint8_t dyn[];
    __shared__ uint64_t redu[16];
    __shared__ int32_t bcast[16];
    // the sublists' prefix sums sit where the planes go: built before each
    // chunk of 256 entries is fetched, overwritten by the planes after
    int *pre = (int *)dyn;  // [groups + 1]
    __shared__ int red[16], s_carry;
    __shared__ uint32_t s_ent[256];  // the workgroup's next entries, fetched together
    const int G = (int)gridDim.x;
    for (int e0 = (int)blockIdx.x;; e0 += 256 * G) {
        __syncthreads();  // the previous entries' planes are no longer read
        if (threadIdx.x == 0) s_carry = 0;
        for (int g0 = 0; g0 < groups; g0 += blockDim.x) {  // exclusive prefix of the counts
            const int g = g0 + (int)threadIdx.x;
            const int n = g < groups ? min(gload(ctl + g), cap) : 0;
            int total;
            const int ex = block_scan_excl(n, red, &total);
            if (g < groups) pre[g] = s_carry + ex;
            __syncthreads();
            if (threadIdx.x == 0) s_carry += total;
            __syncthreads();
        }
        if (threadIdx.x == 0) pre[groups] = s_carry;
        __syncthreads();
        const int count = pre[groups];
        if (e0 >= count) break;
        {
            const int e = e0 + (int)threadIdx.x * G;
            if (e < count) {
                int k = 0;  // the last sublist whose entries start at or before e
                for (int h = STEM_MAX_GROUPS / 2; h > 0; h >>= 1)
                    if (k + h < groups && pre[k + h] <= e) k += h;
                s_ent[threadIdx.x] = gload(list + (size_t)k * cap + (e - pre[k]));
            }
        }
        __syncthreads();
        for (int j = 0; j < 256 && e0 + j * G < count; ++j) {
            const uint32_t ent = s_ent[j];
            const int2 it = items[ent >> 4];
            const int c = (int)(ent & 15u);
            const DevJob &J = jobs[it.x];
            // The descriptor fields first, held in scalar registers; then the
            // generation counter, the planes and the parent's stem from both
            // record buffers, loaded together (only the choice of record
            // waits for the counter).  The child sits at slot N + child of
            // the next buffer, (gen + 1) & 1.
            const SelState *p_st = J.st;
            CdsStat *cs0 = J.cst[0], *cs1 = J.cst[1];
            const uint2 *p_pl = J.planes;
            int32_t *p_sm = J.stem_max;
            int L = J.L, m = J.m, N = J.N;
            asm volatile("" : "+s"(p_st), "+s"(cs0), "+s"(cs1), "+s"(p_pl), "+s"(p_sm), "+s"(L), "+s"(m), "+s"(N));
            const int n = 3 * L, slot = N + it.y, tid = (int)threadIdx.x;
            const size_t rec = (size_t)slot * m + c;
            const int nwords = 2 * scan_words(n);
            const uint64_t *gp = (const uint64_t *)(p_pl + rec * nwords);
            const uint32_t g0 = gload(&p_st->gen);
            const int32_t *r0 = (const int32_t *)(cs0 + rec), *r1 = (const int32_t *)(cs1 + rec);
            const int l0 = gload(r0 + 6), p0 = gload(r0 + 7), q0 = gload(r0 + 8);  // CdsStat::l, p, q
            const int l1 = gload(r1 + 6), p1 = gload(r1 + 7), q1 = gload(r1 + 8);
            const uint64_t v0 = gload(gp + min(tid, nwords - 1));
            uint64_t *planes = (uint64_t *)dyn;  // the CDS's planes as k_design left them
            __syncthreads();  // the previous entry's LDS is no longer read
            if (tid < nwords) planes[tid] = v0;
            for (int i = tid + (int)blockDim.x; i < nwords; i += blockDim.x) planes[i] = gload(gp + i);
            const bool b1 = ((__builtin_amdgcn_readfirstlane(g0) + 1u) & 1u) != 0;
            CdsStat *cs = (b1 ? cs1 : cs0) + rec;
            __syncthreads();
            const uint64_t key = stem_child(n, (const uint2 *)planes, b1 ? l1 : l0, b1 ? p1 : p0, b1 ? q1 : q0, bcast,
                                            redu, (uint16_t *)(planes + nwords));
            if (tid == 0 && key) {
                atomicMax((unsigned long long *)&cs->key, (unsigned long long)key);
                atomicMax(&p_sm[slot], (int)(key >> 40));
            }
        }
    }
}

// Objective 5 of a design: the longest stem over its CDS, ties to the last
// CDS (calMaximumSL's obj5 = l / ((3L-4)/2), utils.cuh:880).  Stores the
// stems k_stem found and clears the marks k_design left.
__device__ int stem_finish_at(CdsStat *cst, float *obj, int8_t *oidx, int32_t *pql, int m, int den_sl, int slot) {
    // a record's 16-byte words 1 (hsc, gc3, l, p) and 2 (q, pad, key), four
    // records' loads issued together
    u32x4 *cs = (u32x4 *)(cst + (size_t)slot * m);
    int bl = 0, bp = 0, bq = 0, bc = m - 1, changed = 0;
    for (int c0 = 0; c0 < m; c0 += 4) {
        u32x4 w1[4], w2[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = min(c0 + u, m - 1);
            w1[u] = gload(cs + 4 * c + 1);
            w2[u] = gload(cs + 4 * c + 2);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = c0 + u;
            if (c >= m) break;
            int l = (int)w1[u].z, p = (int)w1[u].w, q = (int)w2[u].x;
            if (w2[u].y) {  // pad: the stem k_stem found
                stem_unkey((uint64_t)w2[u].z | ((uint64_t)w2[u].w << 32), l, p, q);
                w1[u].z = (uint32_t)l;
                w1[u].w = (uint32_t)p;
                w2[u].x = (uint32_t)q;
                w2[u].y = 0u;
                gstore(cs + 4 * c + 1, w1[u]);
                gstore(cs + 4 * c + 2, w2[u]);
                ++changed;
            }
            if (l >= bl) {
                bl = l;
                bp = p;
                bq = q;
                bc = c;
            }
        }
    }
    gstore(obj + (size_t)slot * NOBJ + 5, (float)bl / (float)den_sl);
    gstore(oidx + (size_t)slot * 2 * NOBJ + 10, (int8_t)bc);
    int32_t *pq = pql + (size_t)slot * 4;
    gstore(pq, bp);
    gstore(pq + 1, bq);
    gstore(pq + 2, bl);
    gstore(pq + 3, 0);
    return changed;
}
__device__ int stem_finish(const DevJob &J, int buf, int slot) {
    return stem_finish_at(J.cst[buf], J.obj[buf], J.oidx[buf], J.pql[buf], J.m, J.den_sl, slot);
}

// stem_finish for every work item's design (initialisation and evaluation;
// the offspring generation leaves it to k_dominance).
__global__ __launch_bounds__(256) void k_stem_finish(const DevJob *__restrict__ jobs, const int2 *__restrict__ items,
                                                     int n_items, int mode) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_items) return;
    const int2 it = items[i];
    const DevJob &J = jobs[it.x];
    int buf, slot;
    item_slot(J, mode, it.y, buf, slot);
    stem_finish(J, buf, slot);
}

// ---------------------------------------------------------------------------
// Selection
// ---------------------------------------------------------------------------
// The objectives of the pool, buffer (gen + 1) & 1.  Both buffers' pointers
// are read from the descriptor before the choice and held in scalar
// registers, so that the rows' loads wait for the generation counter only,
// not for a further descriptor load at an offset the counter decides.
__device__ __forceinline__ const float *pool_obj(const DevJob &J, uint32_t gen) {
    const float *o0 = J.obj[0], *o1 = J.obj[1];
    asm volatile("" : "+s"(o0), "+s"(o1));
    return (gen & 1u) ? o0 : o1;
}

// The pool's extrema over rows row0 .. row0 + rows - 1 (updateIdealValue /
// updateWorstValue, sorting.cuh:132-426; fminf / fmaxf from FLT_MAX /
// -FLT_MAX), folded into the job's accumulators with one atomic minimum and
// maximum per objective.  One wave; `vals` holds the rows' objectives.
__device__ void pool_extrema(const DevJob &J, int row0, int rows, const float *vals) {
    const int lane = threadIdx.x & 63, P = J.P;
    float mn[NOBJ], mx[NOBJ];
#pragma unroll
    for (int k = 0; k < NOBJ; ++k) {
        mn[k] = FLT_MAX;
        mx[k] = -FLT_MAX;
    }
    for (int t = lane; t < rows && row0 + t < P; t += 64)
#pragma unroll
        for (int k = 0; k < NOBJ; ++k) {
            const float v = vals[t * NOBJ + k];
            mn[k] = fminf(mn[k], v);
            mx[k] = fmaxf(mx[k], v);
        }
#pragma unroll
    for (int k = 0; k < NOBJ; ++k) {
        mn[k] = wave_min_f32(mn[k]);
        mx[k] = wave_max_f32(mx[k]);
    }
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NOBJ; ++k) {
            atomicMin(&J.acc->lo[k], ord_bits(mn[k]));
            atomicMax(&J.acc->hi[k], ord_bits(mx[k]));
        }
}

// Bit t of word w of row j is set when design 64w + t dominates design j;
// the matrix is stored word-major, dom[w * P + j], so that the front peeling
// reads one word of consecutive rows per wave.  np[j] counts the dominators.
//
// dominate() (sorting.cuh:94-130) compares |a - b| against the precision
// and the sign of a - b;
// with d = a - b per objective, a dominates b exactly when no d exceeds the
// precision and some d is below minus the precision, and b dominates a
// exactly when the reverse holds (b - a is the negation of a - b, bit for
// bit).  One difference per objective therefore decides both directions,
// and the pool is split into 64 x 64 tiles of (row block b, column block w)
// with b <= w only: item = (job, 64b, w, 0).  Lane t holds column design
// 64w + t.  A tile writes the words of block w of its rows (the ballot of
// its column designs against each row) and, off the diagonal, the words of
// block b of its column designs (each lane gathers one bit per row).  ROWS/64
// tiles per workgroup of four waves, ROWS/4 rows of a tile per wave; the
// partial column words of a tile's waves are combined in LDS.  The diagonal
// tile of a row block also finishes its children's stem records (offspring
// generation) and folds its rows into the pool's extrema.
template <int ROWS>
__global__ __launch_bounds__(256) void k_dominance(const DevJob *__restrict__ jobs, const int4 *__restrict__ items,
                                                   int n_items, bool finish, int32_t *stem_ctl, int ctl_n) {
    static_assert(ROWS == 64 || ROWS == 128 || ROWS == 256, "one, two or four tiles per workgroup");
    constexpr int TPW = ROWS / 64;  // tiles per workgroup
    constexpr int G = 4 / TPW;      // waves per tile
    constexpr int WR = 64 / G;      // rows of a tile per wave
    __shared__ float rows[TPW][64 * NOBJ];
    __shared__ uint64_t part[TPW][G][64];
    __shared__ uint64_t amask[4][WR];  // per wave: the word of block wc of each of its rows
    if (stem_ctl && blockIdx.x == 0)  // k_stem_list has read the counts
        for (int g = threadIdx.x; g < ctl_n; g += blockDim.x) gstore(stem_ctl + g, 0);
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), tl = wv / G, sub = wv % G;
    const int idx = (int)blockIdx.x * TPW + tl;
    const bool have = idx < n_items;  // (the last workgroup may hold fewer tiles)
    const int4 it = items[min(idx, n_items - 1)];
    const DevJob &J = jobs[it.x];
    const int P = J.P, N = J.N, r0 = it.y, wc = it.z, c0 = 64 * wc;
    const bool diag = r0 == c0;
    // the generation counter's load issued, then the descriptor fields read
    // (held in scalar registers: no load waits for the counter's value)
    const uint32_t g0 = gload(&J.st->gen);
    float *o0 = J.obj[0], *o1 = J.obj[1];
    const int32_t *smax = J.stem_max;
    int den = J.den_sl;
    asm volatile("" : "+s"(o0), "+s"(o1), "+s"(smax), "+s"(den));
    // (and those of the stem records' completion below)
    CdsStat *cs0 = J.cst[0], *cs1 = J.cst[1];
    int8_t *oi0 = J.oidx[0], *oi1 = J.oidx[1];
    int32_t *pq0 = J.pql[0], *pq1 = J.pql[1];
    int m = J.m;
    asm volatile("" : "+s"(cs0), "+s"(cs1), "+s"(oi0), "+s"(oi1), "+s"(pq0), "+s"(pq1), "+s"(m));
    const uint32_t gen = __builtin_amdgcn_readfirstlane(g0);
    const int buf = (int)((gen + 1u) & 1u);
    float *const obj = buf ? o1 : o0;
    // (a sort alone has no stem maxima: its loads read the objectives
    // instead, so that they need no branch and issue with the rows')
    if (!finish) smax = (const int32_t *)obj;
    const float den_sl = (float)den;
    // In the offspring generation a child's objective 5 is its longest stem
    // over (3L-4)/2, the value stem_finish stores.  The column design's row
    // and the tile row of the lane are loaded at clamped indices, all before
    // any value is used (the first wave of a tile stages the tile rows).
    const int i = c0 + lane, ic = min(i, P - 1), rr = r0 + lane, rc = min(rr, P - 1);
    float cand[NOBJ], wrow[NOBJ];
#pragma unroll
    for (int k = 0; k < NOBJ; ++k) {
        cand[k] = gload(obj + (size_t)ic * NOBJ + k);
        wrow[k] = gload(obj + (size_t)rc * NOBJ + k);
    }
    const int32_t csm = gload(smax + ic), rsm = gload(smax + rc);
    // the diagonal tiles complete the stem records of their children and
    // count the CDS the children recomputed (before the rows are used, so
    // that the records' loads issue while the rows' are in flight; the
    // objective 5 it stores is the value the rows take below)
    if (finish && diag && sub == 0) {
        int changed = 0;
        if (rr >= N && rr < P && have)
            changed = stem_finish_at(buf ? cs1 : cs0, obj, buf ? oi1 : oi0, buf ? pq1 : pq0, m, den, rr);
        changed = wave_sum_i32(changed);
        if (lane == 0 && changed) atomicAdd(&J.st->cds_scanned, (unsigned long long)changed);
    }
    if (finish && i >= N) cand[NOBJ - 1] = (float)csm / den_sl;
    if (finish && rr >= N) wrow[NOBJ - 1] = (float)rsm / den_sl;
    const bool cv = i < P;
    if (!cv)
#pragma unroll
        for (int k = 0; k < NOBJ; ++k) cand[k] = __builtin_inff();
    if (sub == 0) {
        // rows past the pool hold +inf: no design is dominated by them
        // (d = -inf or NaN in every objective), and their words are not
        // stored
        float *const o = rows[tl] + lane * NOBJ;
#pragma unroll
        for (int k = 0; k < NOBJ; ++k) o[k] = rr < P ? wrow[k] : __builtin_inff();
    }
    __syncthreads();
    // and fold their rows into the pool's extrema
    if (diag && have && sub == G - 1) pool_extrema(J, r0, 64, rows[tl]);
    const float *const rw = rows[tl] + sub * WR * NOBJ;
    uint64_t col = 0;  // bit q: row r0 + sub*WR + q dominates design i
#pragma unroll 1
    for (int q0 = 0; q0 < WR; q0 += 16) {  // 16 rows at a time, last row first
        uint32_t c16 = 0;
        uint64_t am[16];  // (wave-uniform: scalar registers)
#pragma unroll
        for (int u = 15; u >= 0; --u) {
            const float *const row = rw + (q0 + u) * NOBJ;
            float hi = cand[0] - row[0], lo = hi;  // (a NaN difference is never significant)
#pragma unroll
            for (int k = 1; k < NOBJ; ++k) {
                const float d = cand[k] - row[k];
                hi = fmaxf(hi, d);
                lo = fminf(lo, d);
            }
            // the lanes whose design has a difference above the precision
            // and those with one below minus the precision (wave-uniform
            // masks); design i dominates the row when only the second holds
            // (the lanes past the pool hold +inf, so the first), and the row
            // dominates design i when only the first holds
            const uint64_t U = __builtin_amdgcn_ballot_w64(hi > F_PRECISION);
            const uint64_t D = __builtin_amdgcn_ballot_w64(lo < -F_PRECISION);
            am[u] = D & ~U;
            c16 = c16 + c16 + (uint32_t)__builtin_amdgcn_inverse_ballot_w64(U & ~D);
        }
        if (lane == 0)
#pragma unroll
            for (int u = 0; u < 16; ++u) amask[wv][q0 + u] = am[u];
        col |= (uint64_t)c16 << q0;
    }
    __builtin_amdgcn_wave_barrier();
    const uint64_t mine = amask[wv][min(lane, WR - 1)];  // lane q < WR: the word of row r0 + sub*WR + q
    const int r = r0 + sub * WR + lane;
    if (have && lane < WR && r < P) {
        gstore(J.dom + (size_t)wc * P + r, mine);
        if (mine) atomicAdd(&J.np[r], (int)__popcll(mine));
    }
    if (G > 1) {
        part[tl][sub][lane] = col << (sub * WR);
        lds_barrier();
        if (sub == 0)
#pragma unroll
            for (int g = 1; g < G; ++g) col |= part[tl][g][lane];
    }
    if (!diag && have && sub == 0 && cv) {
        gstore(J.dom + (size_t)(r0 >> 6) * P + i, col);
        if (col) atomicAdd(&J.np[i], (int)__popcll(col));
    }
}

__device__ __forceinline__ float asf(const float *x, const float *ideal, int axis) {  // sorting.cuh:428-449
    float mx = -FLT_MAX;
#pragma unroll
    for (int k = 0; k < NOBJ; ++k) {
        float t = x[k] - ideal[k];
        if ((double)t < 0.001) t = 0.f;
        const float v = t / (k == axis ? 1.f : F_PRECISION);
        if (v > mx) mx = v;
    }
    return mx;
}

// GaussianElimination and helpers, sorting.cuh:451-599 (E x = 1)
// Every loop is unrolled so that the matrix stays in registers (the first
// row below the pivot with a usable entry is added through a flag, not an
// index).
__device__ void gauss(const float *E, float *sol) {
    float AB[NOBJ][NOBJ + 1];
#pragma unroll
    for (int r = 0; r < NOBJ; ++r) {
#pragma unroll
        for (int c = 0; c < NOBJ; ++c) AB[r][c] = E[r * NOBJ + c];
        AB[r][NOBJ] = 1.0f;
    }
#pragma unroll
    for (int col = 0; col < NOBJ; ++col) {
        if (fabsf(AB[col][col]) <= 1e-6f) {
            bool added = false;
#pragma unroll
            for (int row = col + 1; row < NOBJ; ++row)
                if (!added && fabsf(AB[row][col]) > 1e-6f) {
#pragma unroll
                    for (int c = col; c < NOBJ + 1; ++c) AB[col][c] += AB[row][c];
                    added = true;
                }
        }
#pragma unroll
        for (int row = col + 1; row < NOBJ; ++row) {
            const float f = AB[row][col] / AB[col][col];
#pragma unroll
            for (int c = col; c < NOBJ + 1; ++c) AB[row][c] -= f * AB[col][c];
        }
    }
#pragma unroll
    for (int row = NOBJ - 1; row >= 0; --row) {
        float s = 0.f;
#pragma unroll
        for (int c = row + 1; c < NOBJ; ++c) s += AB[row][c];
        float x = AB[row][NOBJ];
        if (row != NOBJ - 1) x -= s;
        x /= AB[row][row];
        AB[row][NOBJ] = x;
        AB[row][row] = 1.0f;
#pragma unroll
        for (int r = 0; r < row; ++r) AB[r][row] *= x;
    }
#pragma unroll
    for (int r = 0; r < NOBJ; ++r) sol[r] = AB[r][NOBJ];
}

// (v, id) becomes (ov, oid) when that pair is smaller: a smaller value, or
// the same value with a smaller index (a NaN value is never taken).  Written
// without short-circuit evaluation so that it compiles to selects.
__device__ __forceinline__ void argmin_pair(float &v, int &id, float ov, int oid) {
    const bool take = (ov < v) | ((ov == v) & (oid < id));
    v = take ? ov : v;
    id = take ? oid : id;
}

// argmin_pair over a full wave (DPP row shifts and broadcasts, lane 63 read
// back), wave-uniform; lanes outside a shifted or masked row contribute
// (FLT_MAX, 0x7FFFFFFF), the start value of every lane's fold.
template <int CTRL, int ROWS>
__device__ __forceinline__ void argmin_dpp(float &v, int &id) {
    const float ov = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(FLT_MAX), __float_as_int(v), CTRL, ROWS, 0xf, false));
    const int oid = __builtin_amdgcn_update_dpp(0x7FFFFFFF, id, CTRL, ROWS, 0xf, false);
    argmin_pair(v, id, ov, oid);
}
__device__ __forceinline__ void wave_argmin(float &v, int &id) {
    argmin_dpp<0x111, 0xf>(v, id);
    argmin_dpp<0x112, 0xf>(v, id);
    argmin_dpp<0x114, 0xf>(v, id);
    argmin_dpp<0x118, 0xf>(v, id);
    argmin_dpp<0x142, 0xa>(v, id);
    argmin_dpp<0x143, 0xc>(v, id);
    v = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
    id = __builtin_amdgcn_readlane(id, 63);
}

// asf() for all six axes of one design.  The clamped differences t[k] are +0
// or at least 0.001 (or NaN, which no comparison takes), and axis a takes the
// strict running maximum of t[a] and t[k] / F_PRECISION over k != a.  A
// correctly rounded division by a positive constant is nondecreasing, so the
// largest of those quotients is the quotient of the largest t[k], k != a: the
// largest and second-largest differences give every axis with two divisions,
// the same values as dividing each difference.
__device__ __forceinline__ void asf_all(const float *x, const float *ideal, float *out) {
    float t[NOBJ];
    float m1 = -FLT_MAX, m2 = -FLT_MAX;  // largest and second-largest difference
    int k1 = -1;
#pragma unroll
    for (int k = 0; k < NOBJ; ++k) {
        t[k] = x[k] - ideal[k];
        if ((double)t[k] < 0.001) t[k] = 0.f;
        const bool top = t[k] > m1, second = !top && t[k] > m2;
        m2 = top ? m1 : (second ? t[k] : m2);
        m1 = top ? t[k] : m1;
        k1 = top ? k : k1;
    }
    const float q1 = m1 == -FLT_MAX ? -FLT_MAX : m1 / F_PRECISION;
    const float q2 = m2 == -FLT_MAX ? -FLT_MAX : m2 / F_PRECISION;
#pragma unroll
    for (int a = 0; a < NOBJ; ++a) {
        const float q = a == k1 ? q2 : q1;
        out[a] = t[a] > q ? t[a] : q;
    }
}

__device__ __forceinline__ unsigned long long asf_key(float v, int id) {
    return ((unsigned long long)ord_bits(v) << 32) | (uint32_t)id;
}

// Extreme points (findExtremePoints, sorting.cuh:601-683) settled axis by
// axis against the previous ones, by one wave.  The candidate rows are the
// previous extremes (0-5) and the first front's ASF minima of the six axes
// (6-11), with their ASF values for every axis in casf.  Lane t (0-5) keeps
// the row extreme slot t holds; per axis the slots' values and the first
// front's minimum (lane 6) are reduced to the smallest (value, index) pair --
// the serial fold over the slots in order gives the same pair, NaN values
// never being taken -- and the winner's row goes to the axis's slot.  Returns
// the row held by slot `lane` (lanes 0-5).
__device__ int settle_extremes(const float *pv, const int *pid, const float (*casf)[NOBJ], int P) {
    const int lane = threadIdx.x & 63;
    int cur = lane < NOBJ ? lane : 0;
#pragma unroll
    for (int axis = 0; axis < NOBJ; ++axis) {
        const float c = casf[cur][axis];
        const bool slot = lane < NOBJ && !isnan(c), first = lane == NOBJ;
        float v = slot ? c : (first ? pv[axis] : FLT_MAX);
        int id = slot ? P + lane : (first ? pid[axis] : 0x7FFFFFFF);
        // the smallest pair of lanes 0-7 (prefix minima within the first row)
        argmin_dpp<0x111, 0xf>(v, id);
        argmin_dpp<0x112, 0xf>(v, id);
        argmin_dpp<0x114, 0xf>(v, id);
        v = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 7));
        id = __builtin_amdgcn_readlane(id, 7);
        if (id != 0x7FFFFFFF && v < FLT_MAX) {
            const int t = id - P;  // an extreme slot won: the row it holds
            const int held = (t >= 0 && t < NOBJ) ? __builtin_amdgcn_readlane(cur, t) : NOBJ + axis;
            if (lane == axis) cur = held;
        }
    }
    return cur;
}

// GaussianElimination (sorting.cuh:451-599, E x = 1) by one wave: lane r
// (0-5) holds row r of AB (a[6] the right-hand side) in registers and reads
// the pivot row with v_readlane; every element takes the same operations in
// the same order as the row-by-row elimination, so the solution is the same
// to the bit.  Returns x[lane] in lanes 0-5.
__device__ __forceinline__ float lane_of(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ float gauss_wave(const float *E) {
    const int lane = threadIdx.x & 63;
    const bool row_lane = lane < NOBJ;
    const int r = row_lane ? lane : 0;
    float a[NOBJ + 1];
#pragma unroll
    for (int c = 0; c < NOBJ; ++c) a[c] = E[r * NOBJ + c];
    a[NOBJ] = 1.0f;
#pragma unroll
    for (int col = 0; col < NOBJ; ++col) {
        if (fabsf(lane_of(a[col], col)) <= 1e-6f) {  // add the first row below with a usable entry
            const uint64_t cand = __ballot(row_lane && lane > col && fabsf(a[col]) > 1e-6f);
            if (cand) {
                const int rr = (int)__builtin_ctzll(cand);
#pragma unroll
                for (int c = col; c <= NOBJ; ++c) {
                    const float add = lane_of(a[c], rr);
                    if (lane == col) a[c] += add;
                }
            }
        }
        float prow[NOBJ + 1];
#pragma unroll
        for (int c = col; c <= NOBJ; ++c) prow[c] = lane_of(a[c], col);
        if (row_lane && lane > col) {
            const float f = a[col] / prow[col];
#pragma unroll
            for (int c = col; c <= NOBJ; ++c) a[c] -= f * prow[c];
        }
    }
#pragma unroll
    for (int row = NOBJ - 1; row >= 0; --row) {
        float s = 0.f;
#pragma unroll
        for (int cc = row + 1; cc < NOBJ; ++cc) s += lane_of(a[cc], row);
        float x = lane_of(a[NOBJ], row);
        if (row != NOBJ - 1) x -= s;
        x /= lane_of(a[row], row);
        if (lane == row) {
            a[NOBJ] = x;
            a[row] = 1.0f;
        }
        if (lane < row) a[row] *= x;
    }
    return a[NOBJ];
}

// The first front (designs nobody dominates: their dominator counts are
// complete after k_dominance), one design a thread (grid: jobs x chunks of
// FRONT0_THREADS designs of the largest pool): front_of 0 or -1, the
// first-front words of the ranked mask, and the workgroup's share of the
// first front's size, its maximum per objective (updateNadirValue_MNDF,
// sorting.cuh:685-884) and its per-axis ASF minima against the ideal point
// k_dominance folded (findExtremePoints, sorting.cuh:601-683; the smallest
// value, ties to the lowest index), reduced over the workgroup and folded
// into the job's accumulators with one atomic per quantity.
__global__ __launch_bounds__(FRONT0_THREADS) void k_front0(const DevJob *__restrict__ jobs) {
    const DevJob &J = jobs[blockIdx.x];
    // the descriptor fields held in scalar registers, then the generation
    // counter, the ideal point and the dominator count loaded together (the
    // rows' loads wait for the counter alone)
    SelState *st = J.st;
    SelAcc *acc = J.acc;
    const int32_t *np = J.np;
    const float *o0 = J.obj[0], *o1 = J.obj[1];
    int P = J.P;
    asm volatile("" : "+s"(st), "+s"(acc), "+s"(np), "+s"(o0), "+s"(o1), "+s"(P));
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    constexpr int NW = FRONT0_THREADS / 64;
    const int d0 = blockIdx.y * FRONT0_THREADS;
    if (d0 >= P) return;
    const uint32_t g0 = gload(&st->gen);
    const int d = d0 + tid, dc = min(d, P - 1);
    const int npv = gload(np + dc);
    const uint32_t lo = gload(acc->lo + min(tid, NOBJ - 1));
    const uint32_t gen = __builtin_amdgcn_readfirstlane(g0);
    const float *const obj = (gen & 1u) ? o0 : o1;  // buffer (gen + 1) & 1
    float x[NOBJ];
#pragma unroll
    for (int k = 0; k < NOBJ; ++k) x[k] = gload(obj + (size_t)dc * NOBJ + k);
    __shared__ float ideal[NOBJ], red_v[NW * 2 * NOBJ];
    __shared__ int red_i[NW * NOBJ], s_cnt;
    if (tid < NOBJ) ideal[tid] = ord_float(lo);
    if (tid == 0) s_cnt = 0;
    __syncthreads();
    const bool f0 = d < P && npv == 0;
    if (d < P) gstore(J.front_of + d, f0 ? 0 : -1);
    const uint64_t b = __ballot(f0);  // a wave's 64 designs are one word of the mask
    if (lane == 0 && d0 + 64 * wid < P) {
        gstore(J.ranked + (d0 >> 6) + wid, b);
        if (b) atomicAdd(&s_cnt, (int)__popcll(b));
    }
    float fm[NOBJ], bv[NOBJ];
    int bi[NOBJ];
    asf_all(x, ideal, bv);
#pragma unroll
    for (int k = 0; k < NOBJ; ++k) {
        fm[k] = f0 ? x[k] : -FLT_MAX;
        // an ASF value is never NaN: it is at least -FLT_MAX
        bi[k] = f0 ? d : 0x7FFFFFFF;
        bv[k] = f0 ? bv[k] : FLT_MAX;
        fm[k] = wave_max_f32(fm[k]);
        wave_argmin(bv[k], bi[k]);
    }
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NOBJ; ++k) {
            red_v[wid * 2 * NOBJ + k] = fm[k];
            red_v[wid * 2 * NOBJ + NOBJ + k] = bv[k];
            red_i[wid * NOBJ + k] = bi[k];
        }
    __syncthreads();
    if (tid < NOBJ) {
        float m = -FLT_MAX, v = FLT_MAX;
        int id = 0x7FFFFFFF;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            m = fmaxf(m, red_v[w * 2 * NOBJ + tid]);
            argmin_pair(v, id, red_v[w * 2 * NOBJ + NOBJ + tid], red_i[w * NOBJ + tid]);
        }
        atomicMax(&acc->f0hi[tid], ord_bits(m));
        if (id != 0x7FFFFFFF) atomicMin(&acc->asf[tid], asf_key(v, id));
    }
    if (tid == 0 && s_cnt) atomicAdd(&acc->f0, s_cnt);
}

// One workgroup per job, after k_front0: the selection up to the niching.
//  * Further fronts while fewer than N designs are ranked (nonDominatedSorting's
//    peeling, sorting.cuh:1026-1086).
//  * The extreme points settled against the previous ones and the nadir point
//    (updateNadirValue_HYP with the MNDF fallback, sorting.cuh:911-972), each
//    by one wave.
//  * The survivors of the fronts before the last one, and cleared niche
//    arrays for k_associate.
__global__ __launch_bounds__(1024) void k_fronts(const DevJob *__restrict__ jobs) {
    const DevJob &J = jobs[blockIdx.x];
    const int P = J.P, W = J.W, N = J.N, R = J.R;
    const int tid = threadIdx.x, NT = blockDim.x, lane = tid & 63, wid = tid >> 6;
    // the descriptor fields first, held in scalar registers, then every
    // load of the set-up at once
    SelState *st = J.st;
    const SelAcc *acc = J.acc;
    const uint64_t *rk = J.ranked;
    const float *o0 = J.obj[0], *o1 = J.obj[1];
    asm volatile("" : "+s"(st), "+s"(acc), "+s"(rk), "+s"(o0), "+s"(o1));
    const uint32_t g0 = gload(&st->gen);
    __shared__ uint64_t ranked[MAX_W], newm[MAX_W], lastm[MAX_W];
    __shared__ int s_cnt, s_new;
    __shared__ float ext[NOBJ * NOBJ], ideal[NOBJ], worst[NOBJ], f0max[NOBJ];
    __shared__ float pv[NOBJ], crow[2 * NOBJ][NOBJ], casf[2 * NOBJ][NOBJ];
    __shared__ int pid[NOBJ];
    if (tid < NOBJ) {
        ideal[tid] = ord_float(gload(acc->lo + tid));
        worst[tid] = ord_float(gload(acc->hi + tid));
        f0max[tid] = ord_float(gload(acc->f0hi + tid));
        const unsigned long long k = gload(acc->asf + tid);
        pv[tid] = k == ~0ull ? FLT_MAX : ord_float((uint32_t)(k >> 32));
        pid[tid] = k == ~0ull ? 0x7FFFFFFF : (int)(uint32_t)k;
    }
    if (tid < NOBJ * NOBJ) ext[tid] = gload(st->ext + tid);
    if (tid == 0) {
        s_cnt = gload(&acc->f0);
        s_new = 0;
    }
    const float *const obj = (g0 & 1u) ? o0 : o1;  // buffer (gen + 1) & 1
    for (int w = tid; w < W; w += NT) {
        const uint64_t f0w = gload(rk + w);
        ranked[w] = f0w;
        lastm[w] = f0w;
        newm[w] = 0;
    }
    __syncthreads();
    // candidate rows of the extremes (below): the current extreme points and
    // the rows of the first front's ASF minima, loaded now so that the loads
    // overlap the peeling
    float crv = 0.f;
    if (tid < 2 * NOBJ * NOBJ) {
        const int c = tid / NOBJ, k = tid % NOBJ;
        if (c < NOBJ)
            crv = ext[c * NOBJ + k];
        else if (pid[c - NOBJ] < P)
            crv = gload(obj + (size_t)pid[c - NOBJ] * NOBJ + k);
    }
    int f = 0, last_size = s_cnt;
    while (s_cnt < N) {
        ++f;
        // a design joins front f when all its dominators are ranked
        int c = 0;
        for (int r0 = 0; r0 < P; r0 += NT) {  // a wave's 64 designs are one word of the masks
            const int r = r0 + tid;
            bool joins = false;
            if (r < P && !((ranked[r >> 6] >> (r & 63)) & 1ull)) {
                uint64_t any = 0;
                for (int w = 0; w < W; ++w) any |= gload(J.dom + (size_t)w * P + r) & ~ranked[w];
                joins = !any;
            }
            const uint64_t b = __ballot(joins);
            if (joins) gstore(J.front_of + r, f);
            if (lane == 0 && b) newm[r >> 6] = b;
            c += __popcll(b);
        }
        if (lane == 0) atomicAdd(&s_new, c);
        __syncthreads();
        if (s_new == 0) {  // a dominance cycle: every remaining design forms the front
            for (int r = tid; r < P; r += NT)
                if (!((ranked[r >> 6] >> (r & 63)) & 1ull)) {
                    atomicOr((unsigned long long *)&newm[r >> 6], 1ull << (r & 63));
                    J.front_of[r] = f;
                }
            __syncthreads();
            if (tid == 0) s_new = P - s_cnt;
            __syncthreads();
        }
        for (int w = tid; w < W; w += NT) {
            lastm[w] = newm[w];
            ranked[w] |= newm[w];
            newm[w] = 0;
        }
        last_size = s_new;
        __syncthreads();
        if (tid == 0) {
            s_cnt += s_new;
            s_new = 0;
        }
        __syncthreads();
    }
    const int ranked_n = s_cnt;
    const bool n_cut = ranked_n == N;
    // survivors so far: every ranked design when the fronts fill N exactly,
    // else the fronts before the last one
    if (J.selm)
        for (int w = tid; w < W; w += NT) gstore(J.selm + w, n_cut ? ranked[w] : (ranked[w] & ~lastm[w]));
    if (J.niche && !n_cut) {  // k_associate counts into these
        const NicheView g = niche_view(J.niche, R);
        for (int r = tid; r < R; r += NT) {
            g.cr[r] = make_int2(0, 0);
            gstore(g.fill + r, 0);
            gstore(g.nearest + r, ~0ull);
        }
    }
    // the candidate rows of the extremes and their ASF values for every axis
    if (tid < 2 * NOBJ * NOBJ) crow[tid / NOBJ][tid % NOBJ] = crv;
    __syncthreads();
    if (tid < 2 * NOBJ * NOBJ) casf[tid / NOBJ][tid % NOBJ] = asf(crow[tid / NOBJ], ideal, tid % NOBJ);
    __syncthreads();
    if (wid == 0) {
        const int held = settle_extremes(pv, pid, casf, P);
        const int mine = __shfl(held, lane < NOBJ * NOBJ ? lane / NOBJ : 0, 64);
        if (lane < NOBJ * NOBJ) ext[lane] = crow[mine][lane % NOBJ];
    }
    __syncthreads();
    if (wid == 0) {
        const float x = gauss_wave(ext);  // x[lane], lanes 0-5
        float xs[NOBJ];
#pragma unroll
        for (int i = 0; i < NOBJ; ++i) xs[i] = lane_of(x, i);
        // residuals of every extreme and the intercepts (a lane each)
        bool exc = false;
        float icpt = 0.f;
        if (lane < NOBJ) {
            float res = 0.f;
#pragma unroll
            for (int i = 0; i < NOBJ; ++i) res += xs[i] * (ext[lane * NOBJ + i] - ideal[i]);
            if (fabsf(res - 1.f) > F_PRECISION || isnan(res) || isinf(res)) exc = true;
            icpt = 1.f / x;
            if (icpt <= F_PRECISION || isnan(icpt) || isinf(icpt)) exc = true;
        }
        const bool hyp = __ballot(exc) != 0;
        if (lane < NOBJ) {
            st->ideal[lane] = ideal[lane];
            st->worst[lane] = worst[lane];
            st->nadir[lane] = hyp ? (((f0max[lane] - ideal[lane]) > F_PRECISION) ? f0max[lane] : worst[lane])
                                  : ideal[lane] + icpt;
        }
        if (lane < NOBJ * NOBJ) st->ext[lane] = ext[lane];
        if (lane == 0) {
            st->hyp = hyp;
            st->ranked = ranked_n;
            st->last = f;
            st->base = (f == 0) ? 0 : ranked_n - last_size;
            st->n_cut = n_cut;
        }
    }
}

__device__ __forceinline__ float perp_dist(const float *f, const float *w) {  // utils.cuh:1149-1171
    float num = 0.f, den = 0.f;
#pragma unroll
    for (int i = 0; i < NOBJ; ++i) {
        num += w[i] * f[i];
        den += w[i] * w[i];
    }
    const float k = num / den;
    float d = 0.f;
#pragma unroll
    for (int i = 0; i < NOBJ; ++i) {
        const float t = k * w[i] - f[i];
        d += t * t;
    }
    return sqrtf(d);
}

__device__ __forceinline__ uint64_t niche_prio(const DevJob &J, uint32_t gen, int d) {
    uint32_t r[4];
    philox(gen, (uint32_t)d, 0u, TAG_NICHE, J.seed, r);
    return ((uint64_t)r[0] << 32) | (uint32_t)d;
}

// Nearest reference line of every ranked design (sorting.cuh:1105-1180):
// LPD lanes per design (a wave, or eight lanes), D designs a workgroup (item
// = (job, first design)); the lanes of a design take reference lines
// r = sub, sub + LPD, ... and keep the smallest distance, ties to the lowest
// index (the first minimum of a sweep in index order).  The niche arrays are
// folded in with atomics: designs of the fronts before the last count toward
// their niche, those of the last front toward its bucket, with the bucket's
// nearest design as a 64-bit (distance, design) minimum.  A distance is never
// NaN or above FLT_MAX: the sweep starts at FLT_MAX and takes only smaller
// values.
template <int D, int LPD>
__global__ __launch_bounds__(D * LPD) void k_associate(const DevJob *__restrict__ jobs,
                                                       const int2 *__restrict__ items) {
    static_assert(LPD == 64 || LPD == 8, "a wave or eight lanes per design");
    const int2 it = items[blockIdx.x];
    const DevJob &J = jobs[it.x];
    const int sub = (int)threadIdx.x % LPD, R = J.R, P = J.P;
    const SelState *st = J.st;
    const float *const refpts = J.refpts;
    // everything the wave reads before its first store, loaded together: the
    // selection state, the design's front and objectives (at a clamped index)
    const int n_cut = gload(&st->n_cut), last = gload(&st->last);
    const uint32_t gen = gload(&st->gen);
    float ideal[NOBJ], nadir[NOBJ];
#pragma unroll
    for (int k = 0; k < NOBJ; ++k) {
        ideal[k] = gload(&st->ideal[k]);
        nadir[k] = gload(&st->nadir[k]);
    }
    const int d = it.y + (int)threadIdx.x / LPD, dc = min(d, P - 1);
    // the row in 8-byte words (rows are 24 bytes), each feeding two objectives,
    // so that the loads stay ahead of the per-objective branches
    const f32x2 *o2 = (const f32x2 *)(pool_obj(J, gen) + (size_t)dc * NOBJ);
    const f32x2 x01 = gload(o2), x23 = gload(o2 + 1), x45 = gload(o2 + 2);
    const int fr = gload(J.front_of + dc);
    const float x[NOBJ] = {x01.x, x01.y, x23.x, x23.y, x45.x, x45.y};
    float f[NOBJ];
#pragma unroll
    for (int k = 0; k < NOBJ; ++k) {
        const float den = nadir[k] - ideal[k];
        f[k] = (den == 0.f) ? 0.f : (x[k] - ideal[k]) / den;
    }
    if (n_cut || d >= P || fr < 0) return;  // no niching, or no ranked design (all the design's lanes)
    float best = FLT_MAX;
    int br = 0;
    for (int r = sub; r < R; r += LPD) {
        const float dd = perp_dist(f, refpts + (size_t)r * NOBJ);
        if (dd < best) {
            best = dd;
            br = r;
        }
    }
    bool writer;
    if (LPD == 64) {
        wave_argmin(best, br);  // (every lane of the wave is here: one design a wave)
        writer = sub == 0;
    } else {
        // the design's eight lanes (two quads of a DPP row): the quads by
        // quad permutations, then the first quad into the second by a row
        // shift, whose last lane ends with the design's minimum (the lanes
        // read only their own design's lanes, which are all here)
        argmin_dpp<0xB1, 0xf>(best, br);   // quad_perm [1,0,3,2]
        argmin_dpp<0x4E, 0xf>(best, br);   // quad_perm [2,3,0,1]
        argmin_dpp<0x114, 0xf>(best, br);  // row_shr:4
        writer = sub == 7;
    }
    if (writer) {
        J.ref_of[d] = br;
        J.dist[d] = best;
        const NicheView g = niche_view(J.niche, R);
        if (fr < last) {
            atomicAdd(&g.cr[br].x, 1);
        } else {
            atomicAdd(&g.cr[br].y, 1);
            atomicMin(g.nearest + br, asf_key(best, d));
            J.prio[d] = (uint32_t)(niche_prio(J, gen, d) >> 32);
        }
    }
}

// The levels of referenceBasedSorting (sorting.cuh:1182-1279) in closed form,
// by wave 0 of a workgroup (every thread calls it; the niche arrays are in
// LDS or global memory).  Picks made through level i: C(i) = sum_r
// clamp(i + 1 - cnt[r], 0, rem[r]), nondecreasing in i; the final level is
// the smallest i with C(i) >= need (or P), found by bisection, where the
// first eligible reference points in index order take one more.  Per
// reference point: its picks, its level-0 pick (an empty niche takes its
// nearest design, ties to the lowest index), and, for a bucket only part of
// whose rest is taken (a cut bucket), its offset in the cut list.  Returns
// the cut list's size.
__device__ __forceinline__ int niche_levels(const DevJob &J, const int2 *cr, const unsigned long long *nearest,
                                            int *kp, int *lvl0, int *coff) {
    const int P = J.P, N = J.N, R = J.R, lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __shared__ int s_ncut;
    if (wid == 0) {
        const int need = N - J.st->base;
        // up to 256 reference points: four a lane, in registers
        const bool small = R <= 256;
        int2 rc[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int r = lane + 64 * k;
            rc[k] = small && r < R ? cr[r] : make_int2(0, 0);
        }
        int lo = -1, clo = 0, hi = P;  // C(lo) < need; C(hi) >= need or hi = P
        while (hi - lo > 1) {
            const int mid = lo + ((hi - lo) >> 1);
            int acc = 0;
            if (small) {
#pragma unroll
                for (int k = 0; k < 4; ++k) acc += min(max(mid + 1 - rc[k].x, 0), rc[k].y);
            } else {
                for (int r = lane; r < R; r += 64) {
                    const int2 c = cr[r];
                    acc += min(max(mid + 1 - c.x, 0), c.y);
                }
            }
            const int C = wave_sum_i32(acc);
            if (C >= need) {
                hi = mid;
            } else {
                lo = mid;
                clo = C;
            }
        }
        const int istar = hi, need_final = need - clo;
        int carry = 0, cut_carry = 0;
        for (int r0 = 0; r0 < R; r0 += 64) {
            const int r = r0 + lane;
            const int2 c = r < R ? cr[r] : make_int2(0, 0);
            const bool e = r < R && c.x <= istar && istar < c.x + c.y;
            const uint64_t eb = __ballot(e);
            const int ex = carry + (int)__popcll(eb & ((1ull << lane) - 1ull));
            const int k = min(max(istar - c.x, 0), c.y) + ((e && ex < need_final) ? 1 : 0);
            // level 0 of an empty niche: its nearest design, ties to the
            // lowest index (the reference's fallback to the lowest index, for a
            // NaN distance or none below FLT_MAX, never applies: every distance
            // is at most FLT_MAX and none is NaN)
            const int l0 = (r < R && k > 0 && c.x == 0) ? (int)(uint32_t)nearest[r] : -1;
            const int h = l0 >= 0 ? 1 : 0;
            const int cut = (r < R && k - h > 0 && k - h < c.y - h) ? c.y : 0;
            const int incl = wave_scan_i32(cut);  // inclusive prefix of the cut sizes over the wave
            if (r < R) {
                kp[r] = k;
                lvl0[r] = l0;
                coff[r] = cut ? cut_carry + incl - cut : -1;
            }
            carry += (int)__popcll(eb);
            cut_carry += __builtin_amdgcn_readlane(incl, 63);
        }
        if (lane == 0) s_ncut = cut_carry;
    }
    __syncthreads();
    return s_ncut;
}

// niche_levels for jobs whose niche arrays do not fit in LDS (R above
// NICHE_STAGE): one workgroup per job, on the arrays in global memory.
__global__ __launch_bounds__(256) void k_levels(const DevJob *__restrict__ jobs) {
    const DevJob &J = jobs[blockIdx.x];
    if (J.st->n_cut || J.R <= NICHE_STAGE) return;
    const NicheView g = niche_view(J.niche, J.R);
    const int ncut = niche_levels(J, g.cr, g.nearest, g.kp, g.lvl0, g.coff);
    if (threadIdx.x == 0) J.acc->ncut = ncut;
}

// Every design of the last front, one a thread (grid: jobs x chunks of NT
// designs of the largest pool).  Each workgroup first settles the levels
// itself (niche_levels on the niche arrays staged in LDS; workgroup 0 also
// keeps them for k_niche) when R fits, else reads what k_levels wrote.  A
// design is selected when it is its niche's level-0 pick or when the rest of
// its bucket is taken whole (one atomic OR per wave and word of the survivor
// mask); the designs of a cut bucket get slots in the cut list, bucket by
// bucket, keyed by (priority, design), for k_niche to rank -- the workgroup
// counts its designs per bucket in LDS and claims each bucket's share with
// one atomic (when R fits).
__global__ __launch_bounds__(1024) void k_place(const DevJob *__restrict__ jobs) {
    const DevJob &J = jobs[blockIdx.x];
    const SelState *st = J.st;
    const int P = J.P, R = J.R, NT = blockDim.x, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int d0 = blockIdx.y * NT;
    if (d0 >= P) return;
    // the design's own fields, issued with the state's at a clamped index
    // (ref_of and prio are read only for designs of the last front)
    const int d = d0 + tid, dc = min(d, P - 1);
    const int fo = gload(J.front_of + dc), rf = gload(J.ref_of + dc);
    const uint32_t pr = gload(J.prio + dc);
    if (st->n_cut) return;
    const NicheView g = niche_view(J.niche, R);
    const bool staged = R <= NICHE_STAGE;
    __shared__ int2 s_cr[NICHE_STAGE];
    __shared__ unsigned long long s_near[NICHE_STAGE];
    __shared__ int s_kp[NICHE_STAGE], s_l0[NICHE_STAGE], s_co[NICHE_STAGE], l_cnt[NICHE_STAGE], l_base[NICHE_STAGE];
    const int *kp = g.kp, *lvl0 = g.lvl0, *coff = g.coff;
    if (staged) {
        for (int r = tid; r < R; r += NT) {
            s_cr[r] = g.cr[r];
            s_near[r] = gload(g.nearest + r);
            l_cnt[r] = 0;
        }
        __syncthreads();
        const int ncut = niche_levels(J, s_cr, s_near, s_kp, s_l0, s_co);
        if (blockIdx.y == 0) {
            for (int r = tid; r < R; r += NT) {
                gstore(g.kp + r, s_kp[r]);
                gstore(g.lvl0 + r, s_l0[r]);
                gstore(g.coff + r, s_co[r]);
            }
            if (tid == 0) J.acc->ncut = ncut;
        }
        kp = s_kp;
        lvl0 = s_l0;
        coff = s_co;
    }
    const bool mine = d < P && fo == st->last;
    int r = 0, co = -1, loc = 0;
    bool sel = false;
    if (mine) {
        r = rf;
        const int l0 = lvl0[r];
        co = coff[r];
        sel = d == l0;
        if (co >= 0)
            loc = staged ? atomicAdd(&l_cnt[r], 1) : 0;
        else if (!sel)  // not cut: the rest of the bucket is taken whole, or nothing of it
            sel = kp[r] - (l0 >= 0 ? 1 : 0) > 0;
    }
    if (staged) {
        __syncthreads();
        for (int q = tid; q < R; q += NT)
            if (l_cnt[q]) l_base[q] = atomicAdd(g.fill + q, l_cnt[q]);
        __syncthreads();
    }
    if (mine && co >= 0) {
        const int slot = co + (staged ? l_base[r] + loc : atomicAdd(g.fill + r, 1));
        gstore(J.cutl + slot, ((uint64_t)pr << 32) | (uint32_t)d);
        gstore(J.cutr + slot, r);
    }
    const uint64_t b = __ballot(sel);  // a wave's 64 designs are one word of the mask
    if (lane == 0 && b) atomicOr((unsigned long long *)J.selm + (d0 >> 6) + wid, b);
}

// The designs of the cut buckets, one a thread in cut-list order (grid: jobs
// x chunks of NT designs of the largest pool): a design other than its
// niche's level-0 pick is selected when fewer than kk of the rest of its
// bucket (the bucket without the level-0 pick) have a smaller (priority,
// design) key.  A workgroup stages the keys of every bucket its slots touch
// in LDS (dynamic, 8 bytes a design of the pool).  Workgroup 0 ends the
// generation (nothing here reads the generation count).
__global__ __launch_bounds__(1024) void k_niche(const DevJob *__restrict__ jobs) {
    extern __shared__ __align__(16) uint8_t dyn[];
    uint64_t *keys = (uint64_t *)dyn;
    const DevJob &J = jobs[blockIdx.x];
    SelState *st = J.st;
    const int R = J.R, NT = blockDim.x, tid = threadIdx.x;
    const int s0 = blockIdx.y * NT;
    if (blockIdx.y == 0 && tid == 0) st->gen = st->gen + 1u;
    if (st->n_cut || s0 >= J.acc->ncut) return;
    const int ncut = J.acc->ncut;
    const NicheView g = niche_view(J.niche, R);
    __shared__ int s_a, s_b;
    const int s = s0 + tid, sc = min(s, ncut - 1);
    const uint64_t my = gload(J.cutl + sc);
    const int r = gload(J.cutr + sc);
    // the bucket's fields, loaded together
    const int co = gload(g.coff + r), nr = gload((const int32_t *)(g.cr + r) + 1), l0 = gload(g.lvl0 + r),
              kr = gload(g.kp + r);
    if (tid == 0) s_a = co;  // the first slot's bucket starts here
    if (s == min(s0 + NT, ncut) - 1) s_b = co + nr;
    __syncthreads();
    const int a = s_a, nk = s_b - s_a;
    for (int u = tid; u < nk; u += NT) keys[u] = gload(J.cutl + a + u);
    __syncthreads();
    if (s < ncut) {
        const int d = (int)(uint32_t)my;
        if (d != l0) {
            const int kk = kr - (l0 >= 0 ? 1 : 0);
            const int off = co - a;
            int rank = 0;
#pragma unroll 4
            for (int u = 0; u < nr; ++u) rank += (int)(keys[off + u] < my);
            // the level-0 pick of the bucket is not part of the rest
            if (l0 >= 0 && ((((uint64_t)gload(J.prio + l0)) << 32) | (uint32_t)l0) < my) --rank;
            if (rank < kk) atomicOr((unsigned long long *)J.selm + (d >> 6), 1ull << (d & 63));
        }
    }
}

// k_place and k_niche in one workgroup per job, for pools of at most
// PLACE_NICHE_POOL designs whose niche arrays fit in LDS (R <= NICHE_STAGE):
// the levels settled in LDS, the designs of the last front placed (designs
// d = tid, tid + NT, ...), the cut buckets' keys (priority, design) and
// buckets kept in LDS (dynamic, 12 bytes a design of the pool) instead of the
// global cut list, and the cut designs ranked within their buckets after one
// barrier instead of a kernel boundary.  The same selection as k_place then
// k_niche: a design is selected when it is its niche's level-0 pick, when
// the rest of its bucket is taken whole, or when fewer than kk of the rest of
// its cut bucket have a smaller key.  Ends the generation, as k_niche does.
__global__ __launch_bounds__(1024) void k_place_niche(const DevJob *__restrict__ jobs) {
    extern __shared__ __align__(16) uint8_t dyn[];
    const DevJob &J = jobs[blockIdx.x];
    // the descriptor fields first, held in scalar registers, then the state,
    // the reference points' counts and the first designs' rows at once
    SelState *st = J.st;
    int32_t *niche = J.niche;
    const int32_t *front_of = J.front_of, *ref_of = J.ref_of;
    const uint32_t *prio = J.prio;
    int P = J.P, R = J.R;
    asm volatile("" : "+s"(st), "+s"(niche), "+s"(front_of), "+s"(ref_of), "+s"(prio), "+s"(P), "+s"(R));
    const int NT = blockDim.x, tid = threadIdx.x, lane = tid & 63;
    const int n_cut = gload(&st->n_cut), last = gload(&st->last);
    int fo = gload(front_of + min(tid, P - 1)), rf = gload(ref_of + min(tid, P - 1));
    uint32_t pr = gload(prio + min(tid, P - 1));
    // and the reference points' counts and nearest designs (first chunk)
    const NicheView g = niche_view(niche, R);
    const int rc0 = min(tid, R - 1);
    const uint64_t cr0 = gload((const uint64_t *)g.cr + rc0);  // int2 (x low, y high)
    const unsigned long long nn0 = gload(g.nearest + rc0);
    if (tid == 0) st->gen = st->gen + 1u;  // (nothing here reads the generation count)
    if (n_cut) return;
    uint64_t *keys = (uint64_t *)dyn;          // [P] keys of the cut buckets' designs, bucket by bucket
    int32_t *kbuck = (int32_t *)(keys + P);    // [P] the bucket of each key
    __shared__ int2 s_cr[NICHE_STAGE];
    __shared__ unsigned long long s_near[NICHE_STAGE];
    __shared__ int s_kp[NICHE_STAGE], s_l0[NICHE_STAGE], s_co[NICHE_STAGE], l_cnt[NICHE_STAGE];
    __shared__ uint32_t s_pl0[NICHE_STAGE];  // priority of each bucket's level-0 pick
    if (tid < R) {
        s_cr[tid] = make_int2((int)(uint32_t)cr0, (int)(uint32_t)(cr0 >> 32));
        s_near[tid] = nn0;
        l_cnt[tid] = 0;
    }
    for (int r = tid + NT; r < R; r += NT) {
        s_cr[r] = g.cr[r];
        s_near[r] = gload(g.nearest + r);
        l_cnt[r] = 0;
    }
    __syncthreads();
    const int ncut = niche_levels(J, s_cr, s_near, s_kp, s_l0, s_co);
    for (int r = tid; r < R; r += NT) {
        gstore(g.kp + r, s_kp[r]);
        gstore(g.lvl0 + r, s_l0[r]);
        gstore(g.coff + r, s_co[r]);
    }
    if (tid == 0) J.acc->ncut = ncut;
    // place the designs of the last front
    for (int d0 = 0; d0 < P; d0 += NT) {
        const int d = d0 + tid;
        if (d0) {  // (the first designs' rows were loaded before the levels)
            const int dc = min(d, P - 1);
            fo = gload(front_of + dc);
            rf = gload(ref_of + dc);
            pr = gload(prio + dc);
        }
        const bool mine = d < P && fo == last;
        bool sel = false;
        if (mine) {
            const int l0 = s_l0[rf], co = s_co[rf];
            sel = d == l0;
            if (sel) s_pl0[rf] = pr;  // (the level-0 pick is a last-front design of its bucket)
            if (co >= 0) {
                const int slot = co + atomicAdd(&l_cnt[rf], 1);
                keys[slot] = ((uint64_t)pr << 32) | (uint32_t)d;
                kbuck[slot] = rf;
            } else if (!sel) {  // not cut: the rest of the bucket is taken whole, or nothing of it
                sel = s_kp[rf] - (l0 >= 0 ? 1 : 0) > 0;
            }
        }
        const uint64_t b = __ballot(sel);  // a wave's 64 designs are one word of the mask
        if (lane == 0 && b) atomicOr((unsigned long long *)J.selm + ((d0 + (tid & ~63)) >> 6), b);
    }
    __syncthreads();  // the cut buckets' keys are complete
    // rank the designs of the cut buckets
    for (int sl = tid; sl < ncut; sl += NT) {
        const uint64_t my = keys[sl];
        const int r = kbuck[sl], d = (int)(uint32_t)my, l0 = s_l0[r];
        if (d == l0) continue;
        const int kk = s_kp[r] - (l0 >= 0 ? 1 : 0), off = s_co[r], nr = s_cr[r].y;
        int rank = 0;
#pragma unroll 4
        for (int u = 0; u < nr; ++u) rank += (int)(keys[off + u] < my);
        // the level-0 pick of the bucket is not part of the rest
        if (l0 >= 0 && ((((uint64_t)s_pl0[r]) << 32) | (uint32_t)l0) < my) --rank;
        if (rank < kk) atomicOr((unsigned long long *)J.selm + (d >> 6), 1ull << (d & 63));
    }
}

// Initial state after k_design<INIT>: the first N designs survive, extreme points
// from the first designs (main.cu:101, 106-113).
__global__ void k_init_state(const DevJob *__restrict__ jobs) {
    const DevJob &J = jobs[blockIdx.x];
    SelState *st = J.st;
    for (int w = threadIdx.x; w < J.W; w += blockDim.x) {
        const int n = J.N - 64 * w;
        J.selm[w] = n >= 64 ? ~0ull : (n > 0 ? (1ull << n) - 1ull : 0ull);
    }
    for (int i = threadIdx.x; i < J.P; i += blockDim.x) J.front_of[i] = -1;
    if (threadIdx.x < NOBJ * NOBJ) {
        const int e = threadIdx.x / NOBJ, k = threadIdx.x % NOBJ;
        st->ext[threadIdx.x] = (e < J.N) ? J.obj[0][(size_t)e * NOBJ + k] : 0.f;
    }
    if (threadIdx.x == 0) {
        st->gen = 0;
        st->n_cut = st->hyp = st->ranked = st->last = st->base = 0;
        st->cds_scanned = 0;
    }
}

// Clear the dominator counts and selection accumulators before a standalone
// sort (the generation loop clears them
// inside k_design<OFFSPRING>).
__global__ void k_zero_np(const DevJob *__restrict__ jobs) {
    const DevJob &J = jobs[blockIdx.x];
    for (int i = threadIdx.x; i < J.P; i += blockDim.x) J.np[i] = 0;
    if (threadIdx.x < ACC_WORDS) ((uint32_t *)J.acc)[threadIdx.x] = acc_clear_word(threadIdx.x);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_design(int mode, const DevJob *jobs, const int2 *items, int n, size_t lds, hipStream_t s,
                         uint32_t *stem_list, int32_t *stem_ctl, StemList sl, int item_base, int threads) {
    if (n <= 0) return hipSuccess;
    switch (mode) {
    case 0: hipLaunchKernelGGL((k_design<0, 256>), dim3(n), dim3(256), lds, s, jobs, items, nullptr, nullptr, 1, 0, 0); break;
    case 1: hipLaunchKernelGGL((k_design<1, 256>), dim3(n), dim3(256), lds, s, jobs, items, nullptr, nullptr, 1, 0, 0); break;
    default:
        if (!stem_list || !stem_ctl || sl.cap <= 0 || sl.group <= 0 || sl.groups > STEM_MAX_GROUPS || item_base < 0)
            return hipErrorInvalidValue;
        if (threads == 256)
            hipLaunchKernelGGL((k_design<2, 256>), dim3(n), dim3(256), lds, s, jobs, items, stem_list, stem_ctl,
                               sl.group, sl.cap, item_base);
        else if (threads == 512)
            hipLaunchKernelGGL((k_design<2, 512>), dim3(n), dim3(512), lds, s, jobs, items, stem_list, stem_ctl,
                               sl.group, sl.cap, item_base);
        else
            return hipErrorInvalidValue;
        break;
    }
    return hipGetLastError();
}
int stem_parts_of(int L) { return stem_parts(3 * L); }

hipError_t launch_stem(int mode, const DevJob *jobs, const int2 *items, int n, int mg, int parts, size_t lds,
                       hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (parts < 1 || parts > STEM_PARTS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_stem, dim3(n * mg * parts), dim3(256), lds, s, jobs, items, mode, mg, n * mg);
    return hipGetLastError();
}
hipError_t launch_stem_list(const DevJob *jobs, const int2 *items, const uint32_t *stem_list,
                            const int32_t *stem_ctl, StemList sl, int waves, size_t lds, hipStream_t s) {
    if (sl.cap <= 0 || sl.groups <= 0) return hipSuccess;
    if (sl.groups > STEM_MAX_GROUPS) return hipErrorInvalidValue;
    const int capacity = sl.groups * sl.cap;
    if (lds < 4 * ((size_t)sl.groups + 1)) lds = 4 * ((size_t)sl.groups + 1);  // the prefix sums share the planes' LDS
    if (!stem_list || !stem_ctl || waves < 5 || waves > 8) return hipErrorInvalidValue;
    // workgroups take the entries round-robin, so every one should be
    // resident from the start: per CU, the waves per SIMD or what the
    // registers and the LDS (dynamic and static) hold, as the runtime counts it
    const void *fn = waves == 5   ? (const void *)k_stem_list<5>
                     : waves == 6 ? (const void *)k_stem_list<6>
                     : waves == 7 ? (const void *)k_stem_list<7>
                                  : (const void *)k_stem_list<8>;
    int fit = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, fn, 256, lds) != hipSuccess || fit < 1) fit = 1;
    const int per_cu = fit < waves ? fit : waves;
    const int grid = capacity < STEM_CUS * per_cu ? capacity : STEM_CUS * per_cu;
    switch (waves) {
    case 5: hipLaunchKernelGGL(k_stem_list<5>, dim3(grid), dim3(256), lds, s, jobs, items, stem_list, stem_ctl, sl.groups, sl.cap); break;
    case 6: hipLaunchKernelGGL(k_stem_list<6>, dim3(grid), dim3(256), lds, s, jobs, items, stem_list, stem_ctl, sl.groups, sl.cap); break;
    case 7: hipLaunchKernelGGL(k_stem_list<7>, dim3(grid), dim3(256), lds, s, jobs, items, stem_list, stem_ctl, sl.groups, sl.cap); break;
    default: hipLaunchKernelGGL(k_stem_list<8>, dim3(grid), dim3(256), lds, s, jobs, items, stem_list, stem_ctl, sl.groups, sl.cap); break;
    }
    return hipGetLastError();
}
hipError_t launch_stem_finish(int mode, const DevJob *jobs, const int2 *items, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_stem_finish, dim3((n + 255) / 256), dim3(256), 0, s, jobs, items, n, mode);
    return hipGetLastError();
}
hipError_t launch_dominance(const DevJob *jobs, const int4 *items, int n, bool finish, int32_t *stem_ctl, int ctl_n, int rows,
                            hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int g = (n + rows / 64 - 1) / (rows / 64);
    switch (rows) {
    case 64: hipLaunchKernelGGL(k_dominance<64>, dim3(g), dim3(256), 0, s, jobs, items, n, finish, stem_ctl, ctl_n); break;
    case 128: hipLaunchKernelGGL(k_dominance<128>, dim3(g), dim3(256), 0, s, jobs, items, n, finish, stem_ctl, ctl_n); break;
    case 256: hipLaunchKernelGGL(k_dominance<256>, dim3(g), dim3(256), 0, s, jobs, items, n, finish, stem_ctl, ctl_n); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
hipError_t launch_fronts(const DevJob *jobs, int njobs, int max_pool, int threads, hipStream_t s) {
    if (threads != 256 && threads != 512 && threads != 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_front0, dim3(njobs, (max_pool + FRONT0_THREADS - 1) / FRONT0_THREADS), dim3(FRONT0_THREADS), 0, s,
                       jobs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fronts, dim3(njobs), dim3(threads), 0, s, jobs);
    return hipGetLastError();
}
hipError_t launch_associate(const DevJob *jobs, const int2 *items, int n, int designs, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    switch (designs) {
    case 4: hipLaunchKernelGGL((k_associate<4, 64>), dim3(n), dim3(256), 0, s, jobs, items); break;
    case 8: hipLaunchKernelGGL((k_associate<8, 64>), dim3(n), dim3(512), 0, s, jobs, items); break;
    case 16: hipLaunchKernelGGL((k_associate<16, 64>), dim3(n), dim3(1024), 0, s, jobs, items); break;
    case 32: hipLaunchKernelGGL((k_associate<32, 8>), dim3(n), dim3(256), 0, s, jobs, items); break;
    case 64: hipLaunchKernelGGL((k_associate<64, 8>), dim3(n), dim3(512), 0, s, jobs, items); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
bool shapes_valid(const Shapes &h) {
    return h.design_split >= 0 && (h.design_threads == 256 || h.design_threads == 512) && h.stem_waves >= 5 &&
           h.stem_waves <= 8 &&
           (h.dom_rows == 64 || h.dom_rows == 128 || h.dom_rows == 256) &&
           (h.fronts_threads == 256 || h.fronts_threads == 512 || h.fronts_threads == 1024) &&
           (h.assoc_designs == 4 || h.assoc_designs == 8 || h.assoc_designs == 16 || h.assoc_designs == 32 ||
            h.assoc_designs == 64) &&
           (h.niche_threads == 256 || h.niche_threads == 512 || h.niche_threads == 1024);
}
size_t niche_lds_bytes(int P) { return (size_t)8 * P; }

// The dynamic-LDS limit of a kernel is an attribute of the process, shared by
// every job, batch and captured graph that launches it, so it only ever
// rises: a job created after a larger one must not lower the limit under the
// larger one's launches.  `held` is the limit set so far for that group of
// kernels; `apply` sets the attribute on each of them.
template <class F>
static hipError_t raise_lds_limit(size_t &held, size_t bytes, F apply) {
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    if (bytes <= held) return hipSuccess;
    const hipError_t e = apply((int)bytes);
    if (e == hipSuccess) held = bytes;
    return e;
}
static size_t g_design_lds = 0, g_stem_lds = 0, g_niche_lds = 0;

hipError_t set_niche_lds_limit(size_t bytes) {
    return raise_lds_limit(g_niche_lds, bytes, [](int b) {
        return hipFuncSetAttribute((const void *)k_niche, hipFuncAttributeMaxDynamicSharedMemorySize, b);
    });
}
hipError_t launch_niche(const DevJob *jobs, int njobs, int max_pool, int max_refs, int threads, hipStream_t s) {
    if (threads != 256 && threads != 512 && threads != 1024) return hipErrorInvalidValue;
    hipError_t e = hipSuccess;
    if (max_refs > NICHE_STAGE) {
        hipLaunchKernelGGL(k_levels, dim3(njobs), dim3(256), 0, s, jobs);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (max_pool <= PLACE_NICHE_POOL && max_refs <= NICHE_STAGE) {  // one workgroup per job
        hipLaunchKernelGGL(k_place_niche, dim3(njobs), dim3(threads), place_niche_lds_bytes(max_pool), s, jobs);
        return hipGetLastError();
    }
    const dim3 grid(njobs, (max_pool + threads - 1) / threads);
    hipLaunchKernelGGL(k_place, grid, dim3(threads), 0, s, jobs);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_niche, grid, dim3(threads), niche_lds_bytes(max_pool), s, jobs);
    return hipGetLastError();
}
hipError_t launch_init_state(const DevJob *jobs, int njobs, hipStream_t s) {
    hipLaunchKernelGGL(k_init_state, dim3(njobs), dim3(256), 0, s, jobs);
    return hipGetLastError();
}
hipError_t launch_zero_np(const DevJob *jobs, int njobs, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_np, dim3(njobs), dim3(256), 0, s, jobs);
    return hipGetLastError();
}

hipError_t set_design_lds_limit(size_t bytes) {
    return raise_lds_limit(g_design_lds, bytes, [](int b) {
        const void *fns[] = {(const void *)k_design<0, 256>, (const void *)k_design<1, 256>,
                             (const void *)k_design<2, 256>, (const void *)k_design<2, 512>};
        for (const void *f : fns) {
            const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, b);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    });
}
hipError_t set_stem_lds_limit(size_t bytes) {
    return raise_lds_limit(g_stem_lds, bytes, [](int b) {
        const void *fns[] = {(const void *)k_stem, (const void *)k_stem_list<5>, (const void *)k_stem_list<6>,
                             (const void *)k_stem_list<7>, (const void *)k_stem_list<8>};
        for (const void *f : fns) {
            const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, b);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    });
}

}  // namespace cd
