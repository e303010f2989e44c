// cd_tile.h -- longest hairpin stem of one CDS, bit-parallel across
// anti-diagonals (calMaximumSL, utils.cuh:712-888).
//
// Pair (i, S-i) of anti-diagonal S matches when s[i] == rc[i + n-1-S], rc
// the reverse complement.  With D = n-1-S, the stems are the runs of matches
// along the diagonals of the s-versus-rc match matrix: for D >= 0 (half A)
// position u = i compares s[u] with rc[u + D]; for D < 0 (half B, e = -D)
// position u = i - e compares rc[u] with s[u + e].  Pair (i, S-i) is valid
// (j - i > 4) exactly for u <= (n-6-D)/2, resp. (n-6-e)/2.
//
// A tile is 32 consecutive diagonals (the bits of a word) over 32 consecutive
// positions (rows).  Per row, the 32 match bits are two funnel shifts of the
// sliding window, the row's own base broadcast to all bits, and three logic
// operations on the 2-bit planes; a doubling chain over rows (windows of 2,
// 4 matches, history in registers) marks the rows at which a window of W
// matches ends: W = 8 once the lane's best length so far K is at least 8,
// 7 at K = 7, else 4 (K is at least 4).
// Diagonals so flagged in a tile are measured exactly from the planes, 64
// rows at a time: every maximal run meeting the tile whose length is at least W,
// extended past the tile as far as it goes.  So every run of length
// >= min(K, 8) is measured, and with K never above the longest stem the
// maximum of the keys is exact including the tie rule (longest, then
// smallest p, then smallest q).  The windows 7 and 8 keep flags rare once K
// nears the longest stem; the windows 5 and 6 would need longer histories
// of the match and 2-window rows (more registers than they save).  A CDS
// without any run of 4 is left to an exact scan (the caller checks the
// maximum).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cd_scan.h"

namespace cd {

constexpr int TILE_MIN_RUN = 4;  // shortest window of the tiled pass

// Diagonal blocks per half: diagonals 0 .. n-6 in blocks of 32.
__host__ __device__ inline int tile_blocks(int n) { return n >= 6 ? ((n - 6) >> 5) + 1 : 0; }

// Tiles of the block whose first diagonal is `base`: rows 0 .. (n-6-base)/2.
__host__ __device__ inline int tile_groups(int n, int base) { return (((n - 6 - base) >> 1) >> 5) + 1; }

// Tiles of the first B blocks of a half (closed form of the sum of
// tile_groups over blocks 0..B-1; c = n-6 >= 0, block b has base 32b).
__host__ __device__ inline int tile_prefix(int n, int B) {
    const int c = n - 6;
    const int E = (B + 1) >> 1, O = B >> 1;  // even and odd blocks below B
    int s = E * ((c >> 6) + 1) - E * (E - 1) / 2;
    if (O > 0) s += O * (((c - 32) >> 6) + 1) - O * (O - 1) / 2;
    return s;
}

// Match bits of diagonal d (0 <= d <= n-6) for rows r .. r+31: bit k is row
// r + k, whose base Y[r + k] pairs with X[r + k + d] (X, Y as in tile_rows).
// Bits past the valid rows are not cleared.  For 0 <= r <= (n-6-d)/2 + 32
// the words read stay below scan_words(n).
__host__ __device__ inline uint32_t tile_match(const uint2 *X, const uint2 *Y, int d, int r) {
    const int wy = r >> 5, sy = r & 31, wx = (r + d) >> 5, sx = (r + d) & 31;
    const uint2 y0 = Y[wy], y1 = Y[wy + 1], x0 = X[wx], x1 = X[wx + 1];
    const uint32_t lo = scan_funnel(y1.x, y0.x, sy) ^ scan_funnel(x1.x, x0.x, sx);
    const uint32_t hi = scan_funnel(y1.y, y0.y, sy) ^ scan_funnel(x1.y, x0.y, sx);
    return ~(lo | hi);
}

// Matches of diagonal d from row r upwards (rows r .. umax).
__host__ __device__ inline int tile_run_up(const uint2 *X, const uint2 *Y, int d, int r, int umax) {
    int cnt = 0;
    while (r <= umax) {
        uint32_t w = tile_match(X, Y, d, r);
        if (umax - r < 31) w &= (2u << (umax - r)) - 1u;
        const int t = w == ~0u ? 32 : __builtin_ctz(~w);
        cnt += t;
        if (t < 32) break;
        r += 32;
    }
    return cnt;
}

// Matches of diagonal d from row e downwards (rows e .. 0; 0 if e < 0).
__host__ __device__ inline int tile_run_down(const uint2 *X, const uint2 *Y, int d, int e) {
    int cnt = 0;
    while (e >= 0) {
        const uint32_t w = e >= 31 ? tile_match(X, Y, d, e - 31) : tile_match(X, Y, d, 0) << (31 - e);
        const int t = w == ~0u ? 32 : __builtin_clz(~w);  // bit 31 is row e; rows below 0 shift in as 0
        cnt += t;
        if (t < 32) break;
        e -= 32;
    }
    return cnt;
}

// Exact measure of diagonal d of half h around rows [ua, uz] (uz - ua < 40):
// every maximal run of matches that meets [ua, uz], extended inside the valid
// rows 0 .. (n-6-d)/2, is offered to best (and K) if at least lmin long
// (4 <= lmin <= 8).  Works on 64 rows of match bits at once: the run through
// ua is followed down the diagonal, and runs starting in (ua, uz] are found
// from their first lmin-window; a run reaching the word's last row is
// followed up the diagonal.
__host__ __device__ inline void tile_verify(const uint2 *X, const uint2 *Y, int n, int h, int d, int ua, int uz,
                                            int lmin, uint64_t &best, int &K) {
    const int hi = n - 6 - d;
    if (hi < 0) return;
    const int umax = hi >> 1;
    if (ua < 0) ua = 0;
    if (uz > umax) uz = umax;
    if (ua > uz) return;
    uint64_t m = (uint64_t)tile_match(X, Y, d, ua) | ((uint64_t)tile_match(X, Y, d, ua + 32) << 32);
    if (umax - ua < 63) m &= (2ull << (umax - ua)) - 1ull;  // rows ua .. umax
    auto offer = [&](int rs, int l) {
        if (l < lmin) return;
        const int re = rs + l - 1;
        const int i = h ? re + d : re;
        const int j = h ? n - 1 - re : n - 1 - re - d;
        const uint64_t key = scan_key(l, i - l + 1, j);
        if (key > best) best = key;
        if (l > K) K = l;
    };
    // length of the run starting at bit b (row ua + b), followed past the word
    auto run_from = [&](int b) -> int {
        const uint64_t r = ~(m >> b);  // zeros shifted in from the top end the count at bit 63
        int len = r ? __builtin_ctzll(r) : 64;
        if (b + len == 64) len += tile_run_up(X, Y, d, ua + 64, umax);
        return len;
    };
    // bits clear up to and including bit e
    auto clear_to = [](uint64_t v, int e) -> uint64_t { return e >= 63 ? 0ull : v & ~((2ull << e) - 1ull); };
    const uint64_t m2 = m & (m >> 1), m4 = m2 & (m2 >> 2);
    uint64_t g = m4 & (m4 >> (lmin - 4));  // bit b: rows ua+b .. ua+b+lmin-1 all match
    const int lim = uz - ua;               // <= 39, so every window tested lies inside the word
    g &= (2ull << lim) - 1ull;
    if (m & 1ull) {  // the run through ua
        const int down = tile_run_down(X, Y, d, ua - 1), len = run_from(0);
        offer(ua - down, down + len);
        g = clear_to(g, len - 1);
    }
    while (g) {  // runs starting at bit b in (0, lim] and at least lmin long
        const int b = __builtin_ctzll(g);
        const int len = run_from(b);
        offer(ua + b, len);
        g = clear_to(g, b + len - 1);
    }
}

// Match-chain history carried from one tile of a block to the next: the
// last row's matches, the last two rows' 2-windows, the last four rows'
// 4-windows (index 0 the most recent row).
struct TileHist {
    uint32_t m1, a2[2], a4[4];
};

__host__ __device__ inline void tile_hist_reset(TileHist &t) {
    t.m1 = 0u;
    t.a2[0] = t.a2[1] = 0u;
    t.a4[0] = t.a4[1] = t.a4[2] = t.a4[3] = 0u;
}

// Rows u0 + J0 .. u0 + 31 of block (h, base): X is the plane read along the
// diagonals (rc for half A, s for half B), Y the plane of the row's own base.
// Sets f[0], f[1], f[2] to the diagonals with a window of 4, 7, 8 matches
// ending in these rows (with ALL false only f[2]) and updates the history.  A 7-window is a 4-window
// and the 4-window ending three rows earlier, an 8-window the one ending
// four rows earlier (one three-input logic operation each per row).
template <int J0, bool ALL = true>
__host__ __device__ inline void tile_rows(const uint2 *X, const uint2 *Y, int base, int u0, TileHist &t,
                                          uint32_t f[3]) {
    static_assert((32 - J0) % 8 == 0, "rows in groups of 8");
    const int w = (u0 + base) >> 5;
    const uint2 x0 = X[w], x1 = X[w + 1], y = Y[u0 >> 5];
    uint32_t m1 = t.m1, a2_1 = t.a2[0], a2_2 = t.a2[1];
    uint32_t a4_1 = t.a4[0], a4_2 = t.a4[1], a4_3 = t.a4[2], a4_4 = t.a4[3];
    uint32_t f4 = 0u, f7 = 0u, f8 = 0u;
    // groups of 8 rows: the history rotates with period 4, so an unrolled
    // group renames registers instead of moving them, and the loop over the
    // groups keeps the kernel at 76 VGPRs (6 waves per SIMD)
#pragma unroll 1
    for (int j0 = J0; j0 < 32; j0 += 8) {
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
            const int j = j0 + jj;
            const uint32_t xl = scan_funnel(x1.x, x0.x, j), xh = scan_funnel(x1.y, x0.y, j);
            const uint32_t yl = scan_bit_mask(y.x, j), yh = scan_bit_mask(y.y, j);
            const uint32_t m = ~((xl ^ yl) | (xh ^ yh));
            const uint32_t a2 = m & m1;
            const uint32_t a4 = a2 & a2_2;
            if (ALL) {
                f4 |= a4;
                f7 |= a4 & a4_3;
            }
            f8 |= a4 & a4_4;
            m1 = m;
            a2_2 = a2_1;
            a2_1 = a2;
            a4_4 = a4_3;
            a4_3 = a4_2;
            a4_2 = a4_1;
            a4_1 = a4;
        }
    }
    t.m1 = m1;
    t.a2[0] = a2_1;
    t.a2[1] = a2_2;
    t.a4[0] = a4_1;
    t.a4[1] = a4_2;
    t.a4[2] = a4_3;
    t.a4[3] = a4_4;
    f[0] = f4;
    f[1] = f7;
    f[2] = f8;
}

// One lane's share of the tiled pass: the flat tile range [t0, t1) over half
// A's tiles then half B's (tile_prefix per half); *b, *g track the current
// block and tile, *t the history.  Processes one tile (index T, which must be
// the lane's next tile) and measures its flagged diagonals.
struct TileCursor {
    int h, b, g;  // half, block, tile within block
    TileHist hist;
};

// Place a cursor on flat tile T (T < 2 * tile_prefix(n, nblk)) and rebuild
// the history of the rows before it.
__host__ __device__ inline void tile_seek(const uint2 *sp, const uint2 *rp, int n, int T, TileCursor &c) {
    const int nblk = tile_blocks(n), GA = tile_prefix(n, nblk);
    c.h = T >= GA ? 1 : 0;
    const int Tl = T - c.h * GA;
    int lo = 0, hi = nblk - 1;
    while (lo < hi) {  // the block holding tile Tl: largest b with prefix(b) <= Tl
        const int mid = (lo + hi + 1) >> 1;
        if (tile_prefix(n, mid) <= Tl)
            lo = mid;
        else
            hi = mid - 1;
    }
    c.b = lo;
    c.g = Tl - tile_prefix(n, lo);
    tile_hist_reset(c.hist);
    if (c.g > 0) {  // the last 8 rows of the previous tile of the block
        const uint2 *X = c.h ? sp : rp, *Y = c.h ? rp : sp;
        uint32_t f[3];
        tile_rows<24, false>(X, Y, 32 * c.b, 32 * (c.g - 1), c.hist, f);
    }
}

// Scan the cursor's tile, measure its flagged diagonals, and advance.
__host__ __device__ inline void tile_step(const uint2 *sp, const uint2 *rp, int n, TileCursor &c, int &K,
                                          uint64_t &best) {
    const uint2 *X = c.h ? sp : rp, *Y = c.h ? rp : sp;
    const int base = 32 * c.b, u0 = 32 * c.g;
    const int W = K >= 8 ? 8 : (K == 7 ? 7 : TILE_MIN_RUN);  // K >= TILE_MIN_RUN
    uint32_t fw[3];
    if (W == 8)  // K is the same in the whole wave here, so this branch is too
        tile_rows<0, false>(X, Y, base, u0, c.hist, fw);
    else
        tile_rows<0, true>(X, Y, base, u0, c.hist, fw);
    uint32_t f = W == 8 ? fw[2] : (W == 7 ? fw[1] : fw[0]);
    while (f) {
        const int k = __builtin_ctz(f);
        f &= f - 1u;
        tile_verify(X, Y, n, c.h, base + k, u0 - W + 1, u0 + 31, W, best, K);
    }
    if (++c.g == tile_groups(n, base)) {  // next block (of the next half after the last)
        c.g = 0;
        tile_hist_reset(c.hist);
        if (++c.b == tile_blocks(n)) {
            c.b = 0;
            c.h += 1;
        }
    }
}

}  // namespace cd
