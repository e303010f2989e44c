// cd_qgram.h -- longest hairpin stem of one CDS when a stem length K >= 7 is
// known or guessed: a q-gram seeded search (calMaximumSL, utils.cuh:712-888).
//
// As in cd_tile.h, stems are runs along the diagonals of the s-versus-rc
// match matrix: s[a+t] == rc[b+t] pairs bases (a+t, n-1-b-t), and the pair is
// valid (j - i > 4) while a + b + 2t <= n - 6.  A valid run of R >= K matches
// starting at (a, b) holds, for a stride st <= K - QG_Q + 1, a position
// a* = st*k in [a, a+st-1] whose QG_Q bases lie inside the run; that q-gram of
// s equals the q-gram of rc at b + (a* - a), and the last pair of the q-gram
// is valid.  So the index holds the q-grams of s at a = 0, st, 2st, ... (one
// chain per 12-bit code), every rc position looks up the chain of its q-gram,
// and every hit whose q-gram is valid is measured exactly along its diagonal
// in both directions (cd_tile.h's run counters): every maximal run of length
// >= K is measured.  With K at most the longest stem of the CDS, the maximum
// of the keys is exact including the tie rule (longest, then smallest p, then
// smallest q).  The work is O(n + hits) against the tiles' O(n^2 / 32).
//
// Codes are exact (2 bits per base), so a hit is a true q-gram match; the
// number of hits per lookup is about (n / st) / 4096 for sequences without
// long repeats.  A lookup loop stops at a hit budget, and the caller then
// scans the tiles instead (a periodic sequence can match itself everywhere).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cd_scan.h"
#include "cd_tile.h"

namespace cd {

constexpr int QG_Q = 6;                       // bases per q-gram
constexpr int QG_CODES = 1 << (2 * QG_Q);     // chains in the index
constexpr int QG_KMIN = QG_Q + 1;             // smallest K searched (stride >= 2)
constexpr uint32_t QG_NIL = 0xFFFFu;          // end of a chain

// Stride of the sampled s positions for a search of runs >= K.
__host__ __device__ inline int qg_stride(int K) { return K - QG_Q + 1; }

// Indexed positions a = 0, st, ... with a + QG_Q <= n (0 if n < QG_Q).
__host__ __device__ inline int qg_index_count(int n, int st) { return n >= QG_Q ? (n - QG_Q) / st + 1 : 0; }

// Most index entries of any search on n bases (stride 2).
__host__ __device__ inline int qg_index_max(int n) { return qg_index_count(n, qg_stride(QG_KMIN)); }

// Code of the QG_Q bases of plane P from position a (a + QG_Q <= n): the low
// bits in bits 0-5, the high bits in bits 6-11.  Reads words a/32 and a/32+1,
// both inside scan_words(n).
__host__ __device__ inline uint32_t qg_code(const uint2 *P, int a) {
    const int w = a >> 5, sh = a & 31;
    const uint2 p0 = P[w], p1 = P[w + 1];
    constexpr uint32_t M = (1u << QG_Q) - 1u;
    return (scan_funnel(p1.x, p0.x, sh) & M) | ((scan_funnel(p1.y, p0.y, sh) & M) << QG_Q);
}

// Insert entry k (s position st*k) at the head of its chain.  head holds
// QG_CODES 16-bit entries, QG_NIL when empty, two to a 32-bit word (the
// exchange is a compare-and-swap of that word); next[k] receives the old
// head.
__host__ __device__ inline void qg_insert(const uint2 *sp, int st, int k, uint16_t *head, uint16_t *next) {
    const uint32_t g = qg_code(sp, st * k);
    uint32_t *w = (uint32_t *)head + (g >> 1);
    const int sh = (int)(g & 1u) * 16;
    const uint32_t keep = ~(0xFFFFu << sh), put = (uint32_t)k << sh;
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t cur = *w, prev;
    do {
        prev = cur;
        cur = atomicCAS(w, prev, (prev & keep) | put);
    } while (cur != prev);
#else
    const uint32_t prev = *w;
    *w = (prev & keep) | put;
#endif
    next[k] = (uint16_t)(prev >> sh);
}

// The maximal run through the match (a, b) (a valid pair), clipped to the
// valid pairs of its diagonal, offered to best if at least K long.  One
// 32-row match word from 8 rows below the hit holds the run unless the run
// reaches an end of the word; the run counters then follow it further.
__host__ __device__ inline void qg_measure(const uint2 *sp, const uint2 *rp, int n, int a, int b, int K,
                                           uint64_t &best) {
    const int h = b >= a ? 0 : 1;  // half of cd_tile.h: A reads rc along the diagonal, B reads s
    const uint2 *X = h ? sp : rp, *Y = h ? rp : sp;
    const int d = h ? a - b : b - a, u = h ? b : a;
    const int umax = (n - 6 - d) >> 1;
    const int r0 = u >= 8 ? u - 8 : 0, k = u - r0;  // the hit is bit k of rows r0 .. r0+31
    uint32_t w = tile_match(X, Y, d, r0);
    if (umax - r0 < 31) w &= (2u << (umax - r0)) - 1u;
    const uint32_t ends = ~(w >> k);  // lowest set bit: the first row above the run
    int up = ends ? __builtin_ctz(ends) : 32;
    int down = k ? __builtin_clz(~(w << (32 - k))) : 0;  // matches from row u-1 down, at most k
    if (k + up == 32) up += tile_run_up(X, Y, d, r0 + 32, umax);
    if (down == k && r0 > 0) down += tile_run_down(X, Y, d, r0 - 1);
    const int l = down + up;
    if (l < K) return;
    const int re = u - down + l - 1;  // row of the innermost pair
    const int i = h ? re + d : re, j = h ? n - 1 - re : n - 1 - re - d;
    const uint64_t key = scan_key(l, i - l + 1, j);
    if (key > best) best = key;
}

// One lane's walk over the lookups of its chunks of 8 rc positions (chunk ch
// covers positions 8ch .. 8ch+7 below xe = n - QG_Q + 1; the lane takes the
// chunks ch0, ch0 + stride, ...).  Each step reads one chain entry (or moves
// to the lane's next chunk) and yields the hit it finds when its q-gram is
// valid, so that a wave's steps stay cheap and uniform and the hits can be
// measured apart from the walk.  A chunk's eight heads are read together when
// the chunk is loaded and kept in registers, so a step reads at most one link.
struct QgLane {
    uint32_t lo, hi;  // plane bits of rc from the chunk's first position
    uint32_t mask;    // positions of the chunk whose chains are still to start
    uint32_t h0, h1, h2, h3;  // heads of the chunk's positions, 2j and 2j+1 in the halves of hj
    uint32_t e;       // current chain entry, QG_NIL between chains
    int j;            // position of the current chain within the chunk
    int ch;           // current chunk; the walk is over once it reaches nch
};

// Code of position j of the lane's chunk.
__host__ __device__ inline uint32_t qg_lane_code(const QgLane &s, int j) {
    constexpr uint32_t M = (1u << QG_Q) - 1u;
    return ((s.lo >> j) & M) | (((s.hi >> j) & M) << QG_Q);
}

// Plane bits of chunk ch (one pair of plane words), the heads of its eight
// positions (read unconditionally, so that the reads issue together: a code
// is always below QG_CODES) and its positions with a chain.
__host__ __device__ inline void qg_lane_load(QgLane &s, const uint2 *rp, const uint16_t *head, int xe) {
    const int x0 = s.ch << 3, w = x0 >> 5;
    const uint2 p0 = rp[w], p1 = rp[w + 1];
    s.lo = scan_funnel(p1.x, p0.x, x0 & 31);
    s.hi = scan_funnel(p1.y, p0.y, x0 & 31);
    uint32_t hv[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) hv[j] = head[qg_lane_code(s, j)];
    uint32_t mask = 0u;
#pragma unroll
    for (int j = 0; j < 8; ++j) mask |= (uint32_t)(x0 + j < xe && hv[j] != QG_NIL) << j;
    s.h0 = hv[0] | (hv[1] << 16);
    s.h1 = hv[2] | (hv[3] << 16);
    s.h2 = hv[4] | (hv[5] << 16);
    s.h3 = hv[6] | (hv[7] << 16);
    s.mask = mask;
}

__host__ __device__ inline void qg_lane_start(QgLane &s, const uint2 *rp, const uint16_t *head, int ch0, int nch,
                                              int xe) {
    s.ch = ch0;
    s.e = QG_NIL;
    s.j = 0;
    s.mask = 0u;
    s.lo = s.hi = 0u;
    s.h0 = s.h1 = s.h2 = s.h3 = 0xFFFFFFFFu;
    if (ch0 < nch) qg_lane_load(s, rp, head, xe);
}

__host__ __device__ inline bool qg_lane_done(const QgLane &s, int nch) { return s.ch >= nch; }

// The lane's current chunk has no chain entry left to read.
__host__ __device__ inline bool qg_chunk_done(const QgLane &s) { return s.e == QG_NIL && s.mask == 0u; }

// The lane's next chunk (ch + stride), loaded when there is one.
__host__ __device__ inline void qg_lane_next(QgLane &s, const uint2 *rp, const uint16_t *head, int nch, int stride,
                                             int xe) {
    s.ch += stride;
    if (s.ch < nch) qg_lane_load(s, rp, head, xe);
}

// One chain entry of the lane's current chunk (not qg_chunk_done): returns
// true with the hit (a, b) when the entry's q-gram is valid.  Each entry
// read costs one unit of budget.
__host__ __device__ inline bool qg_walk_step(QgLane &s, const uint16_t *next, int n, int st, int &budget, int &a,
                                             int &b) {
    if (s.e == QG_NIL) {
        const int j = __builtin_ctz(s.mask);
        s.j = j;
        s.mask &= s.mask - 1u;
        // the head of position j (not QG_NIL: the mask has only positions with a chain)
        const uint32_t lo2 = (j & 2) ? s.h1 : s.h0, hi2 = (j & 2) ? s.h3 : s.h2;
        const uint32_t hw = (j & 4) ? hi2 : lo2;
        s.e = (hw >> (16 * (j & 1))) & 0xFFFFu;
    }
    --budget;
    const uint32_t e = s.e;
    s.e = next[e];
    a = st * (int)e;
    b = (s.ch << 3) + s.j;
    return a + b + 2 * (QG_Q - 1) <= n - 6;
}

// One step of a lane that is not done: the next chunk when the current one
// is walked (returns false), else one chain entry (qg_walk_step).
__host__ __device__ inline bool qg_lane_step(QgLane &s, const uint2 *rp, const uint16_t *head, const uint16_t *next,
                                             int n, int st, int nch, int stride, int &budget, int &a, int &b) {
    if (qg_chunk_done(s)) {
        qg_lane_next(s, rp, head, nch, stride, n - QG_Q + 1);
        return false;
    }
    return qg_walk_step(s, next, n, st, budget, a, b);
}

// A hit packed for the queue of hits to measure (both below 2^16).
__host__ __device__ inline uint32_t qg_pack(int a, int b) { return ((uint32_t)a << 16) | (uint32_t)b; }

// Chain entries one lane may read when it looks up `lookups` positions
// (about 0.1 to 0.3 per lookup are expected without long repeats).
__host__ __device__ inline int qg_budget(int lookups) { return 4 * lookups + 64; }

// Hits per wave measured together: the queue of a workgroup of QG_WAVES
// waves holds QG_WAVES * QG_SEG packed hits.
constexpr int QG_SEG = 256;
constexpr int QG_WAVES = 4;

// LDS of a search on n bases: the chain heads, the links of at most
// qg_index_max(n) entries, the hit queue.
__host__ __device__ inline int qg_link_bytes(int n) { return (2 * qg_index_max(n) + 15) & ~15; }
__host__ __device__ inline int qg_lds_bytes(int n) { return QG_CODES * 2 + qg_link_bytes(n) + QG_WAVES * QG_SEG * 4; }

// First K of the search of a changed CDS of n bases, 0 for none: the seed
// (the length of a stem the CDS is known to have) when it is at least
// QG_KMIN, so that the search is exact; else the guess max(QG_KMIN,
// parent_l - 2), two below the parent's longest stem of this CDS, exact when
// the search finds a stem that long.
__host__ __device__ inline int qg_first_k(int n, int seed, int parent_l) {
    if (n > 0xFFFF || (n - 4) / 2 < QG_KMIN) return 0;  // hits are packed in 16-bit halves
    if (seed >= QG_KMIN) return seed;
    return parent_l - 2 > QG_KMIN ? parent_l - 2 : QG_KMIN;
}

// K of the next search after one at K found no stem of K (a guess): QG_KMIN
// after a larger guess, else 0 (the tiles follow).
__host__ __device__ inline int qg_next_k(int K) { return K > QG_KMIN ? QG_KMIN : 0; }

}  // namespace cd
