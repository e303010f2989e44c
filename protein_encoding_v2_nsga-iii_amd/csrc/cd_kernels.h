// cd_kernels.h -- host-side launchers of the codon-design kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "cd_layout.h"

namespace cd {

enum DesignMode { MODE_INIT = 0, MODE_EVAL = 1, MODE_OFFSPRING = 2 };

constexpr size_t LDS_BYTES = 160 * 1024;  // per workgroup on gfx950
constexpr int ASSOC_DESIGNS = 4;           // designs per k_associate workgroup (one wave each)
constexpr int DOM_ROWS = 64;               // rows per k_dominance workgroup (16 per wave)
constexpr int STEM_PARTS = 4;              // most k_stem workgroups per CDS (contiguous tile ranges)

size_t design_lds_bytes(int L, int m);
size_t stem_lds_bytes(int L);
size_t cds_plane_bytes(int L);  // one CDS's two bit planes in the job's plane buffer
int stem_parts_of(int L);  // k_stem workgroups per CDS of L codons (1 .. STEM_PARTS)
hipError_t set_design_lds_limit(size_t bytes);
hipError_t set_stem_lds_limit(size_t bytes);

// items: (job, design or child)
hipError_t launch_design(int mode, const DevJob *jobs, const int2 *items, int n, size_t lds, hipStream_t s);
// parts workgroups per (item, CDS c < mg), parts the largest stem_parts_of of
// the jobs: stem of each CDS k_design marked
hipError_t launch_stem(int mode, const DevJob *jobs, const int2 *items, int n, int mg, int parts, size_t lds,
                       hipStream_t s);
// one thread per item: objective 5, clears the marks
hipError_t launch_stem_finish(int mode, const DevJob *jobs, const int2 *items, int n, hipStream_t s);
// items: (job, first row, word, 0).  finish: the offspring generation, whose
// children's stem objective (and stem records) k_dominance completes in place
// of launch_stem_finish
hipError_t launch_dominance(const DevJob *jobs, const int4 *items, int n, bool finish, hipStream_t s);
// one workgroup per job
hipError_t launch_fronts(const DevJob *jobs, int njobs, hipStream_t s);
// items: (job, first design)
hipError_t launch_associate(const DevJob *jobs, const int2 *items, int n, hipStream_t s);
// one workgroup per job
// one workgroup per job; lds = the largest niche_lds_bytes of the jobs
size_t niche_lds_bytes(int P, int R);
hipError_t set_niche_lds_limit(size_t bytes);
hipError_t launch_niche(const DevJob *jobs, int njobs, size_t lds, hipStream_t s);
hipError_t launch_init_state(const DevJob *jobs, int njobs, hipStream_t s);
hipError_t launch_zero_np(const DevJob *jobs, int njobs, hipStream_t s);

}  // namespace cd
