// cd_noalloc.hip -- non-dominated sorting that allocates nothing.
//
// cd_nds allocates a job-sized block for its scratch (the domination bit
// matrix, dominator counts, selection state and work list).  Callers that
// cannot allocate device memory (code under graph capture, tight serving
// loops) use cd_nds_noalloc instead: the same kernels (k_dominance, k_fronts)
// over the caller's device buffers, with the scratch in one file-scope device
// array of fixed size that every GPU holds for the life of the process.
// Every call on a GPU shares that GPU's array, so calls that overlap on one
// GPU (different streams, other host threads) race; calls in one stream are
// ordered.  What must start at zero (dominator counts, selection state) is
// cleared by a kernel at the start of every call.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "cd_host.h"

namespace {

constexpr int MAX_POOL = CD_NOALLOC_MAX_POOL;
constexpr int MAX_WORDS = (MAX_POOL + 63) / 64;
constexpr int MAX_TILES = MAX_WORDS * (MAX_WORDS + 1) / 2;  // 64 x 64 blocks (b, w), b <= w

constexpr size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t OFF_JOB = 0;
constexpr size_t OFF_ST = OFF_JOB + a256(sizeof(cd::DevJob));
constexpr size_t OFF_NP = OFF_ST + a256(sizeof(cd::SelState));
constexpr size_t OFF_DOM = OFF_NP + a256((size_t)MAX_POOL * 4);
constexpr size_t OFF_TILES = OFF_DOM + a256((size_t)MAX_POOL * MAX_WORDS * 8);
constexpr size_t OFF_ACC = OFF_TILES + a256((size_t)MAX_TILES * sizeof(int4));
constexpr size_t OFF_RANKED = OFF_ACC + a256(sizeof(cd::SelAcc));
constexpr size_t SCRATCH_BYTES = OFF_RANKED + a256((size_t)MAX_WORDS * 8);
static_assert(sizeof(cd::SelState) % 4 == 0, "selection state is cleared in words");

// The scratch of cd_nds_noalloc on every GPU (about 8.2 MiB).  Host code uses
// the address hipGetSymbolAddress returns for the current device, never this
// name.
static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

// Descriptor, cleared dominator counts, selection state and accumulators,
// and the domination tiles (one workgroup).
__global__ __launch_bounds__(1024) void k_noalloc_setup(cd::DevJob h, cd::DevJob *dj, int4 *tiles, int n_tiles,
                                                        int words) {
    for (int i = threadIdx.x; i < h.P; i += blockDim.x) h.np[i] = 0;
    if (threadIdx.x < cd::ACC_WORDS) ((uint32_t *)h.acc)[threadIdx.x] = cd::acc_clear_word(threadIdx.x);
    for (int t = threadIdx.x; t < n_tiles; t += blockDim.x) {  // row block b holds words - b tiles
        int b = 0, first = 0;
        while (first + words - b <= t) first += words - b++;
        tiles[t] = make_int4(0, 64 * b, b + t - first, 0);
    }
    uint32_t *st = (uint32_t *)h.st;
    for (int i = threadIdx.x; i < (int)(sizeof(cd::SelState) / 4); i += blockDim.x) st[i] = 0;
    if (threadIdx.x == 0) *dj = h;
}

__global__ void k_noalloc_finish(const cd::SelState *st, int32_t *n_fronts) {
    if (threadIdx.x == 0) *n_fronts = st->last + 1;
}

std::mutex g_mu;
constexpr int MAX_DEVICES = 64;
uint8_t *g_scratch[MAX_DEVICES];  // per device: the device address of `scratch`

// The address of the current device's scratch, looked up once per device.
int scratch_of_current_device(uint8_t **out) {
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return cdhost::fail(CD_E_DEVICE, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return cdhost::fail(CD_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = (uint8_t *)p;
    }
    *out = g_scratch[dev];
    return CD_OK;
}

}  // namespace

extern "C" int cd_nds_noalloc(const float *obj, int32_t P, int32_t N, int32_t *front_of, int32_t *n_fronts,
                              void *stream) {
    if (!obj || !front_of || !n_fronts) return cdhost::fail(CD_E_INVALID, "null device buffer");
    if (P < 1 || N < 1 || N > P) return cdhost::fail(CD_E_INVALID, "need 1 <= N <= P");
    if (P > MAX_POOL)
        return cdhost::fail(CD_E_INVALID, "pool of " + std::to_string(P) + " designs exceeds cd_nds_noalloc's maximum of " +
                                              std::to_string(MAX_POOL));
    uint8_t *base = nullptr;
    const int rc = scratch_of_current_device(&base);
    if (rc != CD_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int W = (P + 63) / 64;
    const int n_tiles = W * (W + 1) / 2;
    cd::DevJob h{};
    h.N = N;
    h.P = P;
    h.W = W;
    h.R = 1;
    h.den_sl = 1;
    h.obj[0] = h.obj[1] = const_cast<float *>(obj);
    h.front_of = front_of;
    h.np = (int32_t *)(base + OFF_NP);
    h.dom = (uint64_t *)(base + OFF_DOM);
    h.st = (cd::SelState *)(base + OFF_ST);
    h.acc = (cd::SelAcc *)(base + OFF_ACC);
    h.ranked = (uint64_t *)(base + OFF_RANKED);
    cd::DevJob *dj = (cd::DevJob *)(base + OFF_JOB);
    int4 *tiles = (int4 *)(base + OFF_TILES);
    hipLaunchKernelGGL(k_noalloc_setup, dim3(1), dim3(1024), 0, s, h, dj, tiles, n_tiles, W);
    HIPCHK(hipGetLastError());
    HIPCHK(cd::launch_dominance(dj, tiles, n_tiles, false, nullptr, 0, cd::DOM_ROWS, s));
    HIPCHK(cd::launch_fronts(dj, 1, P, cd::SHAPES_JOB.fronts_threads, s));
    hipLaunchKernelGGL(k_noalloc_finish, dim3(1), dim3(64), 0, s, (const cd::SelState *)h.st, n_fronts);
    HIPCHK(hipGetLastError());
    return CD_OK;
}
