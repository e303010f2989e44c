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
constexpr int QG_MAX_INDEX = 0xFFFF;          // indexed positions must stay below QG_NIL

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
// QG_CODES entries, QG_NIL when empty; next[k] receives the old head.
__host__ __device__ inline void qg_insert(const uint2 *sp, int st, int k, uint32_t *head, uint16_t *next) {
    const uint32_t g = qg_code(sp, st * k);
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t old = atomicExch(head + g, (uint32_t)k);
#else
    const uint32_t old = head[g];
    head[g] = (uint32_t)k;
#endif
    next[k] = (uint16_t)old;
}

// The maximal run through the match (a, b) (a valid pair), clipped to the
// valid pairs of its diagonal, offered to best if at least K long.
__host__ __device__ inline void qg_measure(const uint2 *sp, const uint2 *rp, int n, int a, int b, int K,
                                           uint64_t &best) {
    const int h = b >= a ? 0 : 1;  // half of cd_tile.h: A reads rc along the diagonal, B reads s
    const uint2 *X = h ? sp : rp, *Y = h ? rp : sp;
    const int d = h ? a - b : b - a, u = h ? b : a;
    const int umax = (n - 6 - d) >> 1;
    const int down = tile_run_down(X, Y, d, u - 1);
    const int l = down + tile_run_up(X, Y, d, u, umax);
    if (l < K) return;
    const int re = u - down + l - 1;  // row of the innermost pair
    const int i = h ? re + d : re, j = h ? n - 1 - re : n - 1 - re - d;
    const uint64_t key = scan_key(l, i - l + 1, j);
    if (key > best) best = key;
}

// Look up rc position b (b + QG_Q <= n): measure every hit whose q-gram is
// valid.  Each chain entry visited costs one unit of budget; returns false,
// leaving the lookup unfinished, when the budget runs out.
__host__ __device__ inline bool qg_lookup(const uint2 *sp, const uint2 *rp, int n, int st, int K, int b,
                                          const uint32_t *head, const uint16_t *next, int &budget,
                                          uint64_t &best) {
    uint32_t t = head[qg_code(rp, b)];
    while (t != QG_NIL) {
        if (--budget < 0) return false;
        const int a = st * (int)t;
        if (a + b + 2 * (QG_Q - 1) <= n - 6) qg_measure(sp, rp, n, a, b, K, best);
        t = next[t];
    }
    return true;
}

// Hit budget of one lane that looks up `lookups` positions.
__host__ __device__ inline int qg_budget(int lookups) { return 4 * lookups + 64; }

// K of the search for a changed CDS of n bases, 0 for none: the seed (the
// length of a stem the CDS is known to have) when it is at least QG_KMIN, so
// that the search is exact; else the guess parent_l - 2 (two below the
// parent's longest stem of this CDS) when that is at least QG_KMIN, exact
// only when the search finds a stem that long.
__host__ __device__ inline int qg_search_k(int n, int seed, int parent_l) {
    if (qg_index_max(n) >= QG_MAX_INDEX) return 0;
    if (seed >= QG_KMIN) return seed;
    const int g = parent_l - 2;
    return g >= QG_KMIN ? g : 0;
}

}  // namespace cd
