// cd_scan.h -- longest hairpin stem of one CDS, bit-parallel along each
// anti-diagonal (calMaximumSL, utils.cuh:712-888).
//
// A stem with innermost pair (i, j), j - i > MIN_HAIRPIN, is a run of
// consecutive Watson-Crick pairs (i, j), (i-1, j+1), ...  All pairs of one
// run share the anti-diagonal S = i + j.  With rc the reverse complement of
// the sequence (rc[t] = comp(s[n-1-t])), pair (i, S-i) is complementary
// exactly when s[i] == rc[i + n-1-S], so along anti-diagonal S the match bits
// of 32 consecutive positions i are the equality bits of two 32-base windows,
// one of s and one of rc, both read forwards.  Sequences are held as two bit
// planes (low and high bit of the 2-bit code, A=0 C=1 G=2 U=3), so a 32-cell
// chunk costs four funnel shifts and three logic operations.
//
// Runs are found without walking cells: the doubling chain A2 = M & M<<1,
// A4 = A2 & A2<<2, ... marks the positions that end a window of k matches,
// a window of K (the longest stem this lane knows, at least the seed) is
// assembled from it, and only the run ends carrying such a window are
// measured exactly.  Every run of length >= K is measured, and K never
// exceeds the longest stem of the CDS, so the maximum over all lanes is exact
// including the tie rule (longest, then smallest p, then smallest q).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cd {

// Words of one bit plane for n bases, with two zero words of padding (reads
// run one word past the last base), rounded to an even count so that
// 64-position slabs of the plane build never leave the plane.
__host__ __device__ inline int scan_words(int n) { return (((n + 31) >> 5) + 2 + 1) & ~1; }

// ({hi, lo} >> sh) truncated to 32 bits, 0 <= sh <= 31 (v_alignbit_b32).
__host__ __device__ inline uint32_t scan_funnel(uint32_t hi, uint32_t lo, int sh) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbit(hi, lo, (uint32_t)sh);
#else
    return (uint32_t)(((((uint64_t)hi) << 32) | lo) >> sh);
#endif
}

// Bit j of w broadcast to all 32 bits: 0 or ~0 (v_bfe_i32, one op).
__host__ __device__ inline uint32_t scan_bit_mask(uint32_t w, int j) {
#if defined(__HIP_DEVICE_COMPILE__)
    return (uint32_t)__builtin_amdgcn_sbfe((int)w, (uint32_t)j, 1u);
#else
    return 0u - ((w >> j) & 1u);
#endif
}

__host__ __device__ inline uint64_t scan_key(int l, int p, int q) {
    return ((uint64_t)l << 40) | ((uint64_t)(0xFFFFF - p) << 20) | (uint64_t)(0xFFFFF - q);
}

// Window of Kw = min(K, 15) matches: a = floor(log2 Kw), dw = Kw - 2^a.  A
// run of Kw matches ends at k exactly when windows of 2^a matches end at k
// and at k - dw (the two overlap and cover Kw positions).  The level is
// chosen with masks (sel[i] all ones for i == a) so that the per-chunk
// selection is branch-free; shift and dmask encode the delay by dw.  Capping
// the window at 15 only flags more runs (all of them are measured exactly).
struct ScanWindow {
    uint32_t sel[4];
    uint32_t dmask;  // all ones when dw == 0 (no second window)
    int shift;       // 32 - dw, in [1, 31] when dw > 0
};

__host__ __device__ inline void scan_window(int K, ScanWindow &w) {
    const int Kw = K < 15 ? K : 15;
    const int a = Kw >= 8 ? 3 : Kw >= 4 ? 2 : Kw >= 2 ? 1 : 0;
    const int dw = Kw - (1 << a);
    for (int i = 0; i < 4; ++i) w.sel[i] = (i == a) ? ~0u : 0u;
    w.dmask = dw ? 0u : ~0u;
    w.shift = (32 - dw) & 31;
}

__host__ __device__ inline uint32_t scan_level(const ScanWindow &w, uint32_t m, uint32_t a2, uint32_t a4, uint32_t a8) {
    return (m & w.sel[0]) | (a2 & w.sel[1]) | (a4 & w.sel[2]) | (a8 & w.sel[3]);
}

// Scan anti-diagonal S of a sequence of n bases whose planes are sp (sequence,
// .x low bits, .y high bits) and rp (reverse complement), each of
// scan_words(n) words.  K is the lane's window (at least 1; raised when a
// longer stem is found), best its best key.  Pairs (i, S-i) with
// j - i > min_hairpin are valid.
//
// Per chunk the steady state computes the match bits, the windows of 2, 4, 8
// matches and the K-window ends.  Only a chunk with a K-window end (rare once
// K is near the longest stem) measures its runs exactly; a run that reaches
// back past its chunk is followed backwards through earlier chunks, rebuilt
// from the planes.
__host__ __device__ inline void scan_diagonal(const uint2 *sp, const uint2 *rp, int n, int S, int min_hairpin, int &K,
                                              uint64_t &best) {
    const int i0 = S > n - 1 ? S - (n - 1) : 0;
    const int i1 = (S - (min_hairpin + 1)) >> 1;
    if (i1 < i0) return;
    const int len = i1 - i0 + 1;
    const int t0 = i0 + (n - 1 - S);
    const int nch = (len + 31) >> 5;
    const uint32_t lastmask = (len & 31) ? (1u << (len & 31)) - 1u : ~0u;
    const int ss = i0 & 31, st = t0 & 31;
    const uint2 *ps = sp + (i0 >> 5), *pr = rp + (t0 >> 5);
    ScanWindow win;
    scan_window(K, win);
    // match bits of chunk c from words c and c+1 (the last chunk is masked)
    auto chunk_at = [&](int c) -> uint32_t {
        const uint2 a0 = ps[c], a1 = ps[c + 1], b0 = pr[c], b1 = pr[c + 1];
        const uint32_t xl = scan_funnel(a1.x, a0.x, ss) ^ scan_funnel(b1.x, b0.x, st);
        const uint32_t xh = scan_funnel(a1.y, a0.y, ss) ^ scan_funnel(b1.y, b0.y, st);
        return ~(xl | xh) & (c == nch - 1 ? lastmask : ~0u);
    };
    uint32_t Mp = 0, A2p = 0, A4p = 0, A8p = 0, Lp = 0, Wp = 0;
    // runs ending in chunk c-1 (bits Wp) given the next chunk's matches M
    auto measure = [&](uint32_t M, int c) {
        uint32_t E = Wp & ~scan_funnel(M, Mp, 1);
        while (E) {
            const int k = __builtin_ctz(E);
            E &= E - 1u;
            const uint32_t zr = ~Mp & ((2u << k) - 1u);
            int l;
            if (zr) {
                l = k - (31 - __builtin_clz(zr));
            } else {  // the run reaches back past chunk c-1
                l = k + 1;
                for (int cc = c - 2; cc >= 0; --cc) {
                    const uint32_t Mc = chunk_at(cc);
                    if (Mc != ~0u) {
                        l += __builtin_clz(~Mc);
                        break;
                    }
                    l += 32;
                }
            }
            const int ie = i0 + 32 * (c - 1) + k;
            const uint64_t key = scan_key(l, ie - l + 1, S - ie);
            if (key > best) best = key;
            if (l > K) {
                K = l;
                scan_window(K, win);
                Lp = scan_level(win, Mp, A2p, A4p, A8p);
            }
        }
    };
    auto step = [&](uint32_t M, int c) {
        if (Wp) measure(M, c);
        const uint32_t A2 = M & scan_funnel(M, Mp, 31);
        const uint32_t A4 = A2 & scan_funnel(A2, A2p, 30);
        const uint32_t A8 = A4 & scan_funnel(A4, A4p, 28);
        const uint32_t Am = scan_level(win, M, A2, A4, A8);
        Wp = Am & (scan_funnel(Am, Lp, win.shift) | win.dmask);
        Mp = M;
        A2p = A2;
        A4p = A4;
        A8p = A8;
        Lp = Am;
    };
    uint2 slo = ps[0], rlo = pr[0];
    int c = 0;
#pragma unroll 2
    for (; c < nch - 1; ++c) {  // every chunk but the last, streaming the plane words
        const uint2 shi = ps[c + 1], rhi = pr[c + 1];
        const uint32_t xl = scan_funnel(shi.x, slo.x, ss) ^ scan_funnel(rhi.x, rlo.x, st);
        const uint32_t xh = scan_funnel(shi.y, slo.y, ss) ^ scan_funnel(rhi.y, rlo.y, st);
        slo = shi;
        rlo = rhi;
        step(~(xl | xh), c);
    }
    step(chunk_at(c), c);
    if (Wp) measure(0u, c + 1);
}

}  // namespace cd
