// cd_device.h -- device-side primitives of the codon-design engine (gfx950,
// wave64).  Included by the kernel translation unit only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cd_layout.h"

namespace cd {

__constant__ int8_t c_syn_start[21] = {0, 4, 6, 8, 10, 12, 16, 18, 21, 23, 29, 30, 32, 36, 38, 44, 50, 54, 58, 59, 61};
__constant__ int8_t c_syn_num[21] = {4, 2, 2, 2, 2, 4, 2, 3, 2, 6, 1, 2, 4, 2, 6, 6, 4, 4, 1, 2, 3};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint64_t seed,
                                       uint32_t out[4]) {
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        // each 32x32 product whole (one v_mad_u64_u32 instead of a low and a
        // high multiply)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

__device__ __forceinline__ uint32_t uniform_int(uint32_t x, uint32_t k) {
    return (uint32_t)(((uint64_t)x * k) >> 32);
}
__device__ __forceinline__ bool bern(uint32_t x, uint64_t thr) { return (uint64_t)x < thr; }
__host__ __device__ inline uint64_t bern_threshold(float p) {
    double t = (double)p * 4294967296.0;
    if (!(t > 0.0)) return 0;
    if (t >= 18446744073709551615.0) return ~0ull;
    return (uint64_t)t;
}

// exp(x): x = k ln2 + r, Taylor series of exp(r) to degree 14 in Horner form
// with fused multiply-adds.  Only IEEE double operations, so host and device
// produce the same bits.
struct ExpTable {
    double c[15];
};
constexpr ExpTable make_exp_table() {
    ExpTable t{};
    double f = 1.0;
    for (int j = 0; j < 15; ++j) {
        if (j > 0) f *= (double)j;
        t.c[j] = 1.0 / f;
    }
    return t;
}
__constant__ ExpTable c_exp = make_exp_table();

__device__ inline double det_exp(double x) {
    const double INV_LN2 = 0x1.71547652b82fep0;
    const double LN2_HI = 0x1.62e42fefa39efp-1;
    const double LN2_LO = 0x1.abc9e3b39803fp-56;
    double k = __builtin_rint(x * INV_LN2);
    double r = __builtin_fma(-k, LN2_HI, x);
    r = __builtin_fma(-k, LN2_LO, r);
    double p = c_exp.c[14];
#pragma unroll
    for (int j = 13; j >= 0; --j) p = __builtin_fma(p, r, c_exp.c[j]);
    return __builtin_ldexp(p, (int)k);
}

// Stem key: larger is better (longer stem, then smaller p, then smaller q).
__device__ __forceinline__ uint64_t stem_key(int l, int p, int q) {
    return ((uint64_t)l << 40) | ((uint64_t)(0xFFFFF - p) << 20) | (uint64_t)(0xFFFFF - q);
}
__device__ __forceinline__ void stem_unkey(uint64_t k, int &l, int &p, int &q) {
    l = (int)(k >> 40);
    if (l == 0) {
        p = q = 0;
    } else {
        p = 0xFFFFF - (int)((k >> 20) & 0xFFFFF);
        q = 0xFFFFF - (int)(k & 0xFFFFF);
    }
}

// Global-memory access through the global address space.  Pointers held in
// the job descriptor are generic, and generic (flat) accesses also count
// against the LDS wait counter, so an LDS-only barrier would still wait for
// them; these compile to global loads and stores.
template <typename T>
__device__ __forceinline__ T gload(const T *p) {
    return *(const __attribute__((address_space(1))) T *)p;
}
template <typename T>
__device__ __forceinline__ void gstore(T *p, T v) {
    *(__attribute__((address_space(1))) T *)p = v;
}

// 16-byte words (a native 4 x u32 vector, which the global address space
// takes as it is; HIP's vector class does not bind through that cast).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// Workgroup barrier that orders LDS only.  __syncthreads also waits for the
// workgroup's outstanding global stores; code whose threads never read what
// other threads of the workgroup stored to global memory can let those
// stores drain while it goes on.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---------------------------------------------------------------------------
// Wave64 / workgroup reductions
// ---------------------------------------------------------------------------
// Sums over a full wave, wave-uniform: DPP row shifts 1, 2, 4, 8 give each
// lane the sum of its row up to itself (lane 15 the row's sum), the row
// broadcasts of lanes 15 and 31 carry the sums across rows, and lane 63 ends
// with the wave's sum.  Every lane of the wave must be active.
template <int CTRL, int ROWS, bool BOUND>
__device__ __forceinline__ int64_t dpp_i64(int64_t v) {
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, ROWS, 0xf, BOUND);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)((uint64_t)v >> 32), CTRL, ROWS, 0xf, BOUND);
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
// Inclusive prefix sum over a full wave (the steps of wave_sum_i32 without
// the final broadcast).  Every lane of the wave must be active.
__device__ __forceinline__ int wave_scan_i32(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);
    return v;
}
__device__ __forceinline__ int wave_sum_i32(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);
    return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
    v += dpp_i64<0x111, 0xf, true>(v);
    v += dpp_i64<0x112, 0xf, true>(v);
    v += dpp_i64<0x114, 0xf, true>(v);
    v += dpp_i64<0x118, 0xf, true>(v);
    v += dpp_i64<0x142, 0xa, false>(v);
    v += dpp_i64<0x143, 0xc, false>(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), 63);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Smallest and largest v over a full wave, wave-uniform (lanes outside a
// shifted row or a masked row contribute the identity).
__device__ __forceinline__ int wave_min_i32(int v) {
    constexpr int I = 0x7FFFFFFF;
    v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x111, 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x112, 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x114, 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x118, 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x142, 0xa, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x143, 0xc, 0xf, false));
    return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_max_i32(int v) {
    constexpr int I = (int)0x80000000;
    v = max(v, __builtin_amdgcn_update_dpp(I, v, 0x111, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(I, v, 0x112, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(I, v, 0x114, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(I, v, 0x118, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(I, v, 0x142, 0xa, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(I, v, 0x143, 0xc, 0xf, false));
    return __builtin_amdgcn_readlane(v, 63);
}

// Smallest and largest float over a full wave (fminf / fmaxf), wave-uniform;
// the identity is FLT_MAX / -FLT_MAX, the start value of every lane's fold.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_f32(float v, float ident) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(ident), __float_as_int(v), CTRL, ROWS, 0xf, false));
}
__device__ __forceinline__ float wave_min_f32(float v) {
    constexpr float I = 3.402823466e+38f;
    v = fminf(v, dpp_f32<0x111, 0xf>(v, I));
    v = fminf(v, dpp_f32<0x112, 0xf>(v, I));
    v = fminf(v, dpp_f32<0x114, 0xf>(v, I));
    v = fminf(v, dpp_f32<0x118, 0xf>(v, I));
    v = fminf(v, dpp_f32<0x142, 0xa>(v, I));
    v = fminf(v, dpp_f32<0x143, 0xc>(v, I));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ float wave_max_f32(float v) {
    constexpr float I = -3.402823466e+38f;
    v = fmaxf(v, dpp_f32<0x111, 0xf>(v, I));
    v = fmaxf(v, dpp_f32<0x112, 0xf>(v, I));
    v = fmaxf(v, dpp_f32<0x114, 0xf>(v, I));
    v = fmaxf(v, dpp_f32<0x118, 0xf>(v, I));
    v = fmaxf(v, dpp_f32<0x142, 0xa>(v, I));
    v = fmaxf(v, dpp_f32<0x143, 0xc>(v, I));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Largest v >= 0 over a full wave, wave-uniform: DPP row shifts 1, 2, 4, 8
// leave each row's maximum in its lane 15, the row broadcasts of lanes 15
// and 31 carry it across rows, and lane 63 ends with the wave's maximum.
// Every lane of the wave must be active.
__device__ __forceinline__ int wave_max_nonneg(int v) {
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true));   // row_shr:1
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true));   // row_shr:2
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true));   // row_shr:4
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true));   // row_shr:8
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return __builtin_amdgcn_readlane(v, 63);
}

// Largest 64-bit key over a full wave, wave-uniform, by the same DPP steps
// (0, the identity, enters from outside a shifted or masked row: every key is
// at least 0).  Every lane of the wave must be active.
template <int CTRL, int ROWS, bool BOUND>
__device__ __forceinline__ uint64_t dpp_max_u64(uint64_t v) {
    const uint64_t o = (uint64_t)dpp_i64<CTRL, ROWS, BOUND>((int64_t)v);
    return o > v ? o : v;
}
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
    v = dpp_max_u64<0x111, 0xf, true>(v);
    v = dpp_max_u64<0x112, 0xf, true>(v);
    v = dpp_max_u64<0x114, 0xf, true>(v);
    v = dpp_max_u64<0x118, 0xf, true>(v);
    v = dpp_max_u64<0x142, 0xa, false>(v);
    v = dpp_max_u64<0x143, 0xc, false>(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63);
    return ((uint64_t)hi << 32) | lo;
}

__device__ inline uint64_t block_max_u64(uint64_t v, uint64_t *red) {
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    v = wave_max_u64(v);
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    uint64_t m = 0;
    for (int w = 0; w < nw; ++w) m = red[w] > m ? red[w] : m;
    __syncthreads();
    return m;
}

// Exclusive prefix sum of one int per thread over the workgroup; returns the
// prefix and sets *total.  red holds blockDim/64 ints.
__device__ inline int block_scan_excl(int v, int *red, int *total) {
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const int incl = wave_scan_i32(v);  // (every thread calls this)
    __syncthreads();
    if (lane == 63) red[wid] = incl;
    __syncthreads();
    int before = 0, tot = 0;
    for (int w = 0; w < nw; ++w) {
        if (w < wid) before += red[w];
        tot += red[w];
    }
    __syncthreads();
    *total = tot;
    return before + incl - v;
}

}  // namespace cd
