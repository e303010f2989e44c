"""Benchmark of the codon-design engine on one MI355X.

Workload (default): a single NSGA-III job with a synthetic 500-residue protein
(L = 501 codons with the stop), m = 4 CDS copies, population N = 1024
(pool 2048), Das-Dennis H = 5 (252 reference points), Pm = 0.1, seed 2.
A step is one generation: N children mutated and fully evaluated, then
NSGA-III selection of N survivors from the 2N pool.  Inputs and state are
resident in HBM before the timed region; timing uses HIP events on the job's
stream around K graph-launched generations, after W warmup generations.

Prints one JSON line with designs/s (N*K / time), per-generation latency
percentiles, the roofline of the dominant kernel (k_stem, the stem scan of the
changed CDS: VALU-bound, stem cells per second against the int32 VALU peak at
32 cells per lane-op) and a CPU baseline (the oracle on the host cores,
bounded sample).
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
PKG = "protein_encoding_v2_nsga-iii_amd"

# MI355X: 256 CUs x 64 int32 lanes per clock x 2.4 GHz.  A stem cell is one
# (innermost pair, anti-diagonal) test; a bit-sliced scan covers 32 cells per
# 32-bit lane operation, so 32 cells per int32 lane-op is the speed of light
# the scan is measured against.
VALU_INT32_OPS = 256 * 64 * 2.4e9
PEAK_CELLS = 32 * VALU_INT32_OPS
KERNELS = ("offspring", "stem", "stem_finish", "dominance", "fronts", "associate", "niche")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stem_cells(L):
    n = 3 * L
    return (n - 5) * (n - 4) // 2


def workload(args):
    cdp = importlib.import_module(PKG)
    T = cdp.tables
    aa = T.synth_protein(args.residues, args.seed)
    tab = T.synth_tables(args.seed)
    return cdp, aa, tab


def cpu_baseline(aa, tab, args):
    import oracle
    threads = min(os.cpu_count() or 1, 16)
    oracle.set_threads(threads)
    job = oracle.Job(aa, tab, args.m, args.N, args.H, args.Pm, args.seed)
    job.init()
    gens = args.cpu_gens
    t0 = time.perf_counter()
    job.step(gens)
    dt = time.perf_counter() - t0
    return {"value": args.N * gens / dt, "unit": "designs/s", "cores": threads, "kind": "port",
            "sample": f"same workload, {gens} generations after initialisation on the CPU oracle "
                      f"(OpenMP over children, {threads} threads), {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--residues", type=int, default=500)
    ap.add_argument("--m", type=int, default=4)
    ap.add_argument("--N", type=int, default=1024)
    ap.add_argument("--H", type=int, default=5)
    ap.add_argument("--Pm", type=float, default=0.1)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-gens", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB three times on the GPU (a plain copy "
                         "kernel of known size, to calibrate hardware byte counters); not part of the measurement")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(jobs shard across GPUs with no communication, see DESIGN.md)"}))
        sys.exit(2)

    cdp, aa, tab = workload(args)
    L = len(aa)
    job = cdp.Job(aa, tab, args.m, args.N, args.H, args.Pm, args.seed).init()
    log(f"job created and initialised (L={L}, m={args.m}, N={args.N}, H={args.H})")
    job.step(args.warmup)
    job.sync()
    log(f"{args.warmup} warmup generations done")
    ms = job.step_timed(args.steps)
    log(f"{args.steps} timed generations: {ms:.1f} ms")
    designs_s = args.N * args.steps / (ms / 1e3)
    # per-generation latency
    lat = np.array([job.step_timed(1) for _ in range(min(100, args.steps))])
    # dominant kernel duration (HIP events around each launch on the job's stream)
    c0 = job.cds_scanned()
    prof = [float(x) for x in job.profile(20)]
    cds_scanned = (job.cds_scanned() - c0) / 20.0  # CDS scanned per stem launch
    cells = cds_scanned * stem_cells(L)
    achieved = cells / (prof[1] / 1e3)
    log(f"per-kernel profile: {prof}")
    out = {
        "metric": "designs/s + p50/p99 latency, batched vs per-instance on MI355X at equal HV (±1%)",
        "value": round(designs_s, 1),
        "unit": "designs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 codons, int32/int64 integer objectives, fp32 CAI/CPB",
        "data": "synthetic",
        "config": {"workload": f"single job: synthetic {args.residues} aa protein (L={L} with stop), m={args.m}, "
                               f"N={args.N} (pool {2 * args.N}), H={args.H}, Pm={args.Pm}, seed {args.seed}",
                   "L": L, "m": args.m, "N": args.N, "H": args.H, "parallelism": "single GPU, per-instance loop"},
        "latency_ms_per_generation": {"p50": round(float(np.percentile(lat, 50)), 4),
                                      "p99": round(float(np.percentile(lat, 99)), 4)},
        "kernel_ms": dict(zip(KERNELS, [round(x, 4) for x in prof])),
        "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 4), "peak": round(PEAK_CELLS / 1e12, 2),
                     "unit": "T stem cells/s (peak: 32 cells per int32 lane-op)",
                     "frac": round(achieved / PEAK_CELLS, 5), "traffic": None, "kernel": "k_stem",
                     "cells_per_launch": int(cells)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(aa, tab, args)
        log(f"cpu baseline: {out['cpu_baseline']['value']:.1f} designs/s")
    job.close()
    print(json.dumps(out), flush=True)
    if args.copy_mib > 0:
        copy_calibration(args.copy_mib)


def copy_calibration(mib):
    """Three device-to-device copies of `mib` MiB (HIP runs them as its copy
    kernel): a kernel of known traffic for scaling hardware byte counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(mib << 20)
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(src), n) or hip.hipMalloc(ctypes.byref(dst), n):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, n)
    for _ in range(3):
        if hip.hipMemcpy(dst, src, n, 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"copied {mib} MiB three times")


if __name__ == "__main__":
    main()
