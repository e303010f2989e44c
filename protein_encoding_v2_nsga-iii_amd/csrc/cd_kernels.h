// cd_kernels.h -- host-side launchers of the codon-design kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "cd_layout.h"

namespace cd {

enum DesignMode { MODE_INIT = 0, MODE_EVAL = 1, MODE_OFFSPRING = 2 };

constexpr size_t LDS_BYTES = 160 * 1024;  // per workgroup on gfx950
constexpr int DOM_ROWS = 64;               // tile rows per k_dominance workgroup of the standalone sort and selection
constexpr int ASSOC_DESIGNS = 4;           // designs per k_associate workgroup (a wave each), the smallest shape
constexpr int STEM_PARTS = 4;              // most k_stem workgroups per CDS (contiguous tile ranges)
constexpr int STEM_CUS = 256;              // compute units of an MI355X: k_stem_list runs a resident grid

// Launch shapes of one generation's kernels.  Every shape gives the same
// results -- a shape only changes how the work is split over workgroups and
// lanes -- so the autotuner may pick any of them while a job runs.
struct Shapes {
    int design_split;    // k_design LDS bytes up to which children launch apart from the rest (0: one launch)
    int design_threads;  // k_design workgroup of the other children (a lone job: all): 256 or 512 threads
    int stem_waves;      // k_stem_list waves per SIMD (workgroups per CU): 5, 6, 7 or 8
    int dom_rows;        // k_dominance tile rows per workgroup (dom_rows / 64 tiles of 64 x 64): 64, 128 or 256
    int fronts_threads;  // k_fronts workgroup: 256, 512 or 1024 threads
    int assoc_designs;   // designs per k_associate workgroup: 4, 8 or 16 (a wave each), 32 or 64 (eight lanes each)
    int niche_threads;   // niching workgroup (k_place_niche, or k_place and k_niche above 2048 designs): 256, 512 or 1024 threads
};
// Untuned shapes: a lone job, and a batch or served tick.
constexpr Shapes SHAPES_JOB = {0, 512, 6, 64, 512, 4, 1024};
constexpr Shapes SHAPES_BATCH = {24 * 1024, 512, 7, 128, 1024, 4, 1024};
bool shapes_valid(const Shapes &s);

size_t design_lds_bytes(int L, int m);
size_t stem_lds_bytes(int L);
size_t cds_plane_bytes(int L);  // one CDS's two bit planes in the job's plane buffer
int stem_parts_of(int L);  // k_stem workgroups per CDS of L codons (1 .. STEM_PARTS)
hipError_t set_design_lds_limit(size_t bytes);
hipError_t set_stem_lds_limit(size_t bytes);

// The changed CDS of an offspring generation, in sublists of consecutive
// children (`group` of them, at least 64, at most STEM_MAX_GROUPS sublists),
// each with its own count in stem_ctl, so that the children's appends spread
// over many counters and the sublists in order keep the children list's
// order (longest jobs first); a sublist holds `cap` entries.
constexpr int STEM_MAX_GROUPS = 2048;
struct StemList {
    int group, groups, cap;
};
inline StemList stem_list_geom(int n_items, int mg) {
    StemList g;
    g.group = (n_items + STEM_MAX_GROUPS - 1) / STEM_MAX_GROUPS;
    if (g.group < 64) g.group = 64;
    g.groups = (n_items + g.group - 1) / g.group;
    g.cap = g.group * mg;
    return g;
}
// k_stem_list's dynamic LDS holds the sublists' prefix sums (at most this
// many bytes) before it holds a CDS's planes and tables
constexpr size_t STEM_PRE_BYTES = 4 * (STEM_MAX_GROUPS + 1);
inline size_t stem_list_lds(size_t stem_lds) { return stem_lds > STEM_PRE_BYTES ? stem_lds : STEM_PRE_BYTES; }
inline size_t stem_list_entries(int n_items, int mg) {
    const StemList g = stem_list_geom(n_items, mg);
    return (size_t)g.groups * g.cap;
}
// items: (job, design or child).  The offspring mode lists every changed CDS
// in the sublists (entry ((item_base + item) << 4) | c; sl =
// stem_list_geom of the whole children list) and counts them in stem_ctl;
// the three are required there.  item_base: the position of items[0] in the
// children list k_stem_list reads.  threads: the offspring mode's workgroup,
// 256 or 512 (initialisation and evaluation take 256).
hipError_t launch_design(int mode, const DevJob *jobs, const int2 *items, int n, size_t lds, hipStream_t s,
                         uint32_t *stem_list = nullptr, int32_t *stem_ctl = nullptr, StemList sl = StemList{0, 0, 0},
                         int item_base = 0, int threads = 256);
// initialisation and evaluation: parts workgroups per (item, CDS c < mg),
// parts the largest stem_parts_of of the jobs: stem of each CDS k_design marked
hipError_t launch_stem(int mode, const DevJob *jobs, const int2 *items, int n, int mg, int parts, size_t lds,
                       hipStream_t s);
// offspring generations: stem of each CDS the offspring k_design listed
// (items and sl as given to it), on as many workgroups as are resident at
// once (STEM_CUS x the workgroups per CU that `waves`, 5 to 8, and the LDS
// allow), at most the entries there can be; launch_dominance with the same
// stem_ctl and sl.groups clears the counts
hipError_t launch_stem_list(const DevJob *jobs, const int2 *items, const uint32_t *stem_list,
                            const int32_t *stem_ctl, StemList sl, int waves, size_t lds, hipStream_t s);
// one thread per item: objective 5, clears the marks
hipError_t launch_stem_finish(int mode, const DevJob *jobs, const int2 *items, int n, hipStream_t s);
// items: (job, 64b, w, 0), the 64 x 64 tiles b <= w, rows / 64 (rows: 64, 128 or 256) per workgroup.  finish: the offspring generation, whose children's stem
// objective (and stem records) k_dominance completes in place of
// launch_stem_finish; stem_ctl (or null): the ctl_n changed-CDS counts of
// that generation's k_design, cleared for the next one
hipError_t launch_dominance(const DevJob *jobs, const int4 *items, int n, bool finish, int32_t *stem_ctl, int ctl_n,
                            int rows, hipStream_t s);
// k_front0 (grid: jobs x chunks of FRONT0_THREADS designs of the largest
// pool, max_pool), then k_fronts (one workgroup of `threads` (256, 512 or
// 1024) per job)
constexpr int FRONT0_THREADS = 256;
hipError_t launch_fronts(const DevJob *jobs, int njobs, int max_pool, int threads, hipStream_t s);
// items: (job, first design), `designs` designs per item: 4, 8 or 16 (a wave each), 32 or 64 (eight lanes each)
hipError_t launch_associate(const DevJob *jobs, const int2 *items, int n, int designs, hipStream_t s);
// Pools of at most PLACE_NICHE_POOL designs with at most NICHE_STAGE
// reference points: k_place_niche (one workgroup of `threads` per job,
// place_niche_lds_bytes(max_pool) of dynamic LDS).  Otherwise k_levels (one
// workgroup per job) when some job has more reference points than
// NICHE_STAGE (max_refs), then k_place and k_niche (grid: jobs x chunks of
// `threads` (256, 512 or 1024) designs of the largest pool, max_pool; k_niche
// takes niche_lds_bytes(max_pool) of dynamic LDS, set_niche_lds_limit above
// 48 KiB)
constexpr int PLACE_NICHE_POOL = 2048;
inline size_t place_niche_lds_bytes(int P) { return (size_t)12 * P; }
size_t niche_lds_bytes(int P);
hipError_t set_niche_lds_limit(size_t bytes);
hipError_t launch_niche(const DevJob *jobs, int njobs, int max_pool, int max_refs, int threads, hipStream_t s);
hipError_t launch_init_state(const DevJob *jobs, int njobs, hipStream_t s);
hipError_t launch_zero_np(const DevJob *jobs, int njobs, hipStream_t s);

}  // namespace cd
