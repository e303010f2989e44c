// cd_stem.h -- bit-sliced scan for the longest hairpin stem of one CDS
// (calMaximumSL, utils.cuh:712-888).
//
// The stem length of an innermost pair (i, j), j - i > 4, is the number of
// consecutive Watson-Crick pairs (i-t, j+t).  With r the reverse complement of
// s (r[t] = comp(s[n-1-t])), pair (i, j) is complementary exactly when
// s[i] == r[n-1-j], so stems are runs on the diagonals of the s-versus-r match
// matrix restricted to i + t <= n-6 (t = n-1-j).
//
// A lane owns one group: 32 consecutive diagonals (the bits of a word) over 32
// consecutive positions u.  Half A covers diagonals d = t - i >= 0 with
// u = i, window X = r at u+d, Y = s[u]; half B covers e = i - t >= 1 with
// u = t, window X = s at u+e, Y = r[u].  Per position the 32 mismatch bits are
// two funnel shifts and three logic operations; a sliding OR over TP positions
// (doubling chain) flags every run of at least TP matches, and the rare flagged
// diagonals are re-walked exactly on the nucleotide codes, which also applies
// the validity limits (u <= (n-6-d)/2) the bit-sliced pass ignores.
#pragma once

#include <stdint.h>

namespace cd {

constexpr int STEM_MIN_TP = 4;  // shortest window of the bit-sliced pass

// Plane layout: word w of a plane holds positions 32w..32w+31; planes are
// padded with STEM_PAD zero words so windows may run past the end.
constexpr int STEM_PAD = 3;

__host__ __device__ inline int stem_words(int n) { return ((n + 31) >> 5) + STEM_PAD; }
__host__ __device__ inline int stem_blocks(int n) { return n >= 6 ? ((n - 6) >> 5) + 1 : 0; }
__host__ __device__ inline int stem_umax(int n, int base) { return (n - 6 - base) >> 1; }
__host__ __device__ inline int stem_groups_of(int n, int base) { return (stem_umax(n, base) >> 5) + 1; }

__host__ __device__ inline uint64_t stem_key_hd(int l, int p, int q) {
    return ((uint64_t)l << 40) | ((uint64_t)(0xFFFFF - p) << 20) | (uint64_t)(0xFFFFF - q);
}

__host__ __device__ inline uint32_t funnel(uint32_t hi, uint32_t lo, int sh) {
    return (uint32_t)(((((uint64_t)hi) << 32) | lo) >> sh);
}

struct StemView {
    const uint8_t *s;      // nucleotide codes, A=0 C=1 G=2 U=3
    const uint32_t *s_lo, *s_hi, *r_lo, *r_hi;
    int n, nblk;           // blocks of 32 diagonals per half
};

// Exact re-walk of diagonal `diag` of half `hb` around positions [ua, uz]:
// every maximal run that meets [ua, uz] is extended to its full length inside
// the valid range and offered to the lane's best key.
__host__ __device__ inline void stem_verify(const StemView &V, int hb, int diag, int ua, int uz, int lmin,
                                            uint64_t &best, int &best_l) {
    const int n = V.n;
    if (hb && diag == 0) return;  // e = 0 is diagonal d = 0 of half A
    const int hi = n - 6 - diag;
    if (hi < 0) return;
    const int umax = hi >> 1;
    if (ua < 0) ua = 0;
    if (uz > umax) uz = umax;
    const uint8_t *s = V.s;
    auto match = [&](int u) -> bool {
        return hb ? ((s[u + diag] ^ s[n - 1 - u]) == 3) : ((s[u] ^ s[n - 1 - u - diag]) == 3);
    };
    int u = ua;
    while (u <= uz) {
        if (!match(u)) {
            ++u;
            continue;
        }
        int rs = u, re = u;
        if (u == ua)
            while (rs > 0 && match(rs - 1)) --rs;
        while (re < umax && match(re + 1)) ++re;
        const int l = re - rs + 1;
        if (l >= lmin) {
            const int i = hb ? re + diag : re;
            const int j = hb ? n - 1 - re : n - 1 - re - diag;
            const uint64_t k = stem_key_hd(l, i - l + 1, j);
            if (k > best) {
                best = k;
                if (l > best_l) best_l = l;
            }
        }
        u = re + 1;
    }
}

// Mismatch-history carried between consecutive groups of one lane.
struct StemHist {
    uint32_t m1;
    uint32_t p2[12];
    uint32_t p4[8];
    uint32_t p8[8];
};

__host__ __device__ inline void stem_hist_reset(StemHist &h) {
    h.m1 = ~0u;
    for (int i = 0; i < 12; ++i) h.p2[i] = ~0u;
    for (int i = 0; i < 8; ++i) h.p4[i] = ~0u;
    for (int i = 0; i < 8; ++i) h.p8[i] = ~0u;
}

// One group: positions u0+J0..u0+31 of the 32 diagonals base..base+31 of
// half hb.  Returns the bits whose TP-window was all-match at some position
// (only when ACC).  h carries the chain across groups.
template <int TP, int J0, bool ACC>
__host__ __device__ inline uint32_t stem_group(const StemView &V, int hb, int base, int u0, StemHist &h) {
    const uint32_t *Xlo = hb ? V.s_lo : V.r_lo, *Xhi = hb ? V.s_hi : V.r_hi;
    const uint32_t *Ylo = hb ? V.r_lo : V.s_lo, *Yhi = hb ? V.r_hi : V.s_hi;
    const int xw = (u0 + base) >> 5, yw = u0 >> 5;
    const uint32_t x0l = Xlo[xw], x1l = Xlo[xw + 1], x0h = Xhi[xw], x1h = Xhi[xw + 1];
    const uint32_t yl = Ylo[yw], yh = Yhi[yw];
    uint32_t M[33], P2[44], P4[40], P8[40];
    M[0] = h.m1;
#pragma unroll
    for (int i = 0; i < 12; ++i) P2[i] = h.p2[i];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        P4[i] = h.p4[i];
        P8[i] = h.p8[i];
    }
    uint32_t acc = 0;
#pragma unroll
    for (int j = J0; j < 32; ++j) {
        const uint32_t xl = funnel(x1l, x0l, j), xh = funnel(x1h, x0h, j);
        const uint32_t ml = 0u - ((yl >> j) & 1u), mh = 0u - ((yh >> j) & 1u);
        const uint32_t m = (xl ^ ml) | (xh ^ mh);
        M[j - J0 + 1] = m;
        const int a = j - J0;
        P2[12 + a] = m | M[a];
        P4[8 + a] = P2[12 + a] | P2[10 + a];
        P8[8 + a] = P4[8 + a] | P4[4 + a];
        uint32_t o;
        if (TP == 4) o = P4[8 + a];
        else if (TP == 6) o = P4[8 + a] | P2[8 + a];
        else if (TP == 8) o = P8[8 + a];
        else if (TP == 10) o = P8[8 + a] | P2[4 + a];
        else if (TP == 12) o = P8[8 + a] | P4[a];
        else if (TP == 14) o = P8[8 + a] | P4[a] | P2[a];
        else o = P8[8 + a] | P8[a];
        if (ACC) acc |= ~o;
    }
    const int cnt = 32 - J0;
    h.m1 = M[cnt];
#pragma unroll
    for (int i = 0; i < 12; ++i) h.p2[i] = P2[cnt + i];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        h.p4[i] = P4[cnt + i];
        h.p8[i] = P8[cnt + i];
    }
    return acc;
}

// Largest instantiated window not above t.
__host__ __device__ inline int stem_tp_for(int t) {
    if (t >= 16) return 16;
    if (t < STEM_MIN_TP) return STEM_MIN_TP;
    return t & ~1;
}

// Groups [G, g1) of one lane with a fixed window TP; returns the first group
// not processed (the window changed or the range ended).  b/g locate group G.
template <int TP, class GetT, class PutT>
__host__ __device__ inline int stem_lane_tp(const StemView &V, const int *gpre, int G, int g1, int &b, int &g,
                                            StemHist &h, uint64_t &best, int &best_l, GetT get_t, PutT put_t) {
    for (; G < g1; ++G) {
        if (stem_tp_for(get_t()) != TP) return G;
        while (G >= gpre[b + 1]) {  // next block starts at u = 0 with no history
            ++b;
            g = 0;
            stem_hist_reset(h);
        }
        const int hb = b >= V.nblk, base = 32 * (hb ? b - V.nblk : b), u0 = 32 * g;
        const uint32_t acc = stem_group<TP, 0, true>(V, hb, base, u0, h);
        if (acc) {
            uint32_t bits = acc;
            while (bits) {
                const int k = __builtin_ctz(bits);
                bits &= bits - 1;
                stem_verify(V, hb, base + k, u0 - TP + 1, u0 + 31, TP, best, best_l);
            }
            if (best_l > get_t()) put_t(best_l);
        }
        ++g;
    }
    return G;
}

// Scan the flat group range [g0, g1) of one lane.  gpre[b] is the number of
// groups before block b (blocks 0..nblk-1 are half A, nblk..2nblk-1 half B).
// get_t() returns the current shared threshold (the longest stem known so
// far, uniform across a wave on the device); put_t(l) publishes a longer one.
template <class GetT, class PutT>
__host__ __device__ inline uint64_t stem_lane(const StemView &V, const int *gpre, int g0, int g1, GetT get_t,
                                              PutT put_t) {
    uint64_t best = 0;
    int best_l = 0;
    if (g0 >= g1) return 0;
    const int nb = 2 * V.nblk;
    int lo = 0, hi = nb - 1;
    while (lo < hi) {  // block holding group g0
        const int mid = (lo + hi + 1) >> 1;
        if (gpre[mid] <= g0)
            lo = mid;
        else
            hi = mid - 1;
    }
    int b = lo, g = g0 - gpre[lo];
    StemHist h;
    stem_hist_reset(h);
    if (g > 0) {  // a range that starts inside a block: rebuild the last 16 positions of history
        const int hb = b >= V.nblk;
        stem_group<16, 16, false>(V, hb, 32 * (hb ? b - V.nblk : b), 32 * (g - 1), h);
    }
    int G = g0;
    while (G < g1) {
        switch (stem_tp_for(get_t())) {
        case 4: G = stem_lane_tp<4>(V, gpre, G, g1, b, g, h, best, best_l, get_t, put_t); break;
        case 6: G = stem_lane_tp<6>(V, gpre, G, g1, b, g, h, best, best_l, get_t, put_t); break;
        case 8: G = stem_lane_tp<8>(V, gpre, G, g1, b, g, h, best, best_l, get_t, put_t); break;
        case 10: G = stem_lane_tp<10>(V, gpre, G, g1, b, g, h, best, best_l, get_t, put_t); break;
        case 12: G = stem_lane_tp<12>(V, gpre, G, g1, b, g, h, best, best_l, get_t, put_t); break;
        case 14: G = stem_lane_tp<14>(V, gpre, G, g1, b, g, h, best, best_l, get_t, put_t); break;
        default: G = stem_lane_tp<16>(V, gpre, G, g1, b, g, h, best, best_l, get_t, put_t); break;
        }
    }
    return best;
}

}  // namespace cd
