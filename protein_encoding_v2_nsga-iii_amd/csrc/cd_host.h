// cd_host.h -- host-side internals shared by the C ABI (cd_api.hip) and the
// serving loop (cd_server.hip): the job and batch objects behind the opaque
// handles, error reporting, and the launch sequence of one generation.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "cd_kernels.h"
#include "codon_nsga3.h"

struct cd_batch;

// One job: tables, pools and selection state in one device allocation.
// The captured generation graphs of a job or batch: [0] one generation,
// [1] GRAPH_BLOCK generations back to back.  A long step launches the block
// graph (consecutive graph launches leave several microseconds between
// them; the kernels inside one graph follow each other closely).
constexpr int GRAPH_BLOCK = 8;
struct StepGraphs {
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    void reset() {
        for (int k = 0; k < 2; ++k) {
            if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
            if (graph[k]) (void)hipGraphDestroy(graph[k]);
            exec[k] = nullptr;
            graph[k] = nullptr;
        }
    }
};

struct cd_job {
    // live batches that hold the job: while any does, the job is advanced
    // only through them (cd_job_step and cd_job_destroy fail) and reads wait
    // for their streams
    std::vector<cd_batch *> batches;
    cd::DevJob h{};
    cd::DevJob *d_self = nullptr;
    int L = 0, m = 0, N = 0, H = 0, R = 0, P = 0, W = 0;
    void *mem = nullptr;
    size_t mem_bytes = 0;
    hipStream_t stream = nullptr;
    int2 *d_off = nullptr;
    int4 *d_dom = nullptr;        // domination tiles: the 64 x 64 blocks (b, w), b <= w
    int2 *d_asc = nullptr;        // association groups, room for the smallest group (4 designs)
    uint32_t *d_slist = nullptr;  // changed CDS of a generation (stem_list_entries) for k_stem_list
    int32_t *d_sctl = nullptr;    // its counts (STEM_MAX_GROUPS), zero between generations (k_dominance clears them)
    int n_dom = 0, n_asc = 0;
    cd::Shapes shapes = cd::SHAPES_JOB;  // launch shapes of its generations (the lists above follow them)
    size_t lds = 0, stem_lds = 0;
    int stem_parts = 1;  // k_stem workgroups per CDS (initialisation and evaluation)
    StepGraphs graphs;
    bool inited = false;
    bool selection_only = false;
};

struct cd_batch {
    std::vector<cd_job *> jobs;
    cd::Shapes shapes = cd::SHAPES_BATCH;  // launch shapes of its ticks (the lists below follow them)
    cd::DevJob *d_jobs = nullptr;
    int2 *d_off = nullptr;
    int4 *d_dom = nullptr;  // the domination tiles of every job
    int2 *d_asc = nullptr;  // room for 4-design groups
    uint32_t *d_slist = nullptr;  // changed CDS of a generation (stem_list_entries)
    int32_t *d_sctl = nullptr;
    int n_off = 0, n_dom = 0, n_asc = 0, n_cds = 0;
    int n_small = 0;        // children first in d_off whose k_design fits six per CU
    size_t lds_small = 0;   // their k_design LDS
    size_t lds = 0, stem_lds = 0;
    int max_pool = 0;    // largest pool of the batch's jobs
    int max_refs = 0;    // most reference points of the batch's jobs
    int mg = 0;          // largest m of the batch's jobs
    int stem_parts = 0;  // largest k_stem workgroups per CDS of the batch's jobs
    hipStream_t stream = nullptr;
    StepGraphs graphs;
};

namespace cdhost {

inline thread_local std::string g_err;  // cd_last_error of the calling thread

inline int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                                          \
    do {                                                                                                      \
        hipError_t e_ = (expr);                                                                               \
        if (e_ != hipSuccess) return ::cdhost::fail(CD_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device work lists and launch sizes of one set of jobs advanced together.
struct Launch {
    const cd::DevJob *jobs;
    int njobs;
    const int2 *off;  // (job, child) of every child
    int n_off;
    const int4 *dom;  // (job, 64b, w, 0) domination tiles, b <= w, taken sh.dom_rows / 64 per workgroup
    int n_dom;
    const int2 *asc;  // (job, first design) association groups of sh.assoc_designs designs
    int n_asc;
    size_t lds, stem_lds;
    int max_pool;    // largest pool
    int max_refs;    // most reference points
    int mg;          // largest m
    int stem_parts;  // largest k_stem workgroups per CDS (initialisation and evaluation)
    uint32_t *stem_list;  // changed CDS of a generation, in sublists of consecutive children (stem_list_entries)
    int32_t *stem_ctl;    // their counts, zero between generations (k_dominance clears them)
    int n_cds;            // CDS of all children: sum of N * m
    // the first n_small children (0: none apart) are of jobs whose k_design
    // LDS, at most lds_small (at most sh.design_split), they are launched
    // apart from the others, which take lds
    int n_small;
    size_t lds_small;
    cd::Shapes sh;
};

// Host images of the work lists of `jobs` (job k = index k of the list)
// and the largest per-job launch sizes.
struct WorkLists {
    std::vector<cd::DevJob> descs;
    std::vector<int2> off, asc;
    std::vector<int4> dom;
    size_t lds = 0, stem_lds = 0;
    int max_pool = 0, max_refs = 0, mg = 0, stem_parts = 0, n_cds = 0;
    int n_small = 0;  // children (first in off) of jobs with lds at most shapes.design_split
    size_t lds_small = 0;
};

// The work lists of a launch sequence over `jobs` with the given shapes.
void build_lists(cd_job *const *jobs, int count, const cd::Shapes &shapes, WorkLists &wl);

// Autotuner.  Shapes picked for a kind of work are cached per device under
// a coarse signature of that work (job count, children, largest codons per
// child, largest pool, each as a power-of-two bucket), so later batches,
// jobs and served ticks of the same kind start from them.
std::string work_signature(cd_job *const *jobs, int count);
bool cached_shapes(const std::string &sig, cd::Shapes *out);
void store_shapes(const std::string &sig, const cd::Shapes &sh);

// Mutation, evaluation and selection of one generation of every job of `w`;
// ev (or null): CD_PROFILE_KERNELS + 1 events recorded around the phases.
hipError_t enqueue_generation(const Launch &w, hipStream_t s, hipEvent_t *ev);

// Raise the dynamic LDS limits of the kernels to what `wl` needs.
hipError_t set_lds_limits(const WorkLists &wl);

}  // namespace cdhost
