"""Benchmark of the codon-design engine on one MI355X.

Workload (default): a single NSGA-III job with a synthetic 500-residue protein
(L = 501 codons with the stop), m = 4 CDS copies, population N = 1024
(pool 2048), Das-Dennis H = 5 (252 reference points), Pm = 0.1, seed 2.
A step is one generation: N children mutated and fully evaluated, then
NSGA-III selection of N survivors from the 2N pool.  Inputs and state are
resident in HBM before the timed region; timing uses HIP events on the job's
stream around K graph-launched generations, after W warmup generations.

Prints one JSON line with designs/s (N*K / time), per-generation latency
percentiles, the roofline of the dominant kernel of the live per-kernel
profile (k_design's offspring launch: HBM-bound, the algorithmic bytes of a
child, parent read plus survivor and child written, against 8 TB/s; the stem
scan (k_stem_list), when it dominates: VALU-bound, stem cells against 32 cells per int32
lane-op), the stem scan's cell rate beside it, and a CPU baseline (the oracle
on the host cores, bounded sample).
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
HBM_BYTES_S = 8.0e12
KERNELS = ("offspring", "stem", "dominance", "fronts", "associate", "niche")
# HBM bytes per launch at the default workload: FETCH_SIZE x 2.000 +
# WRITE_SIZE x 1.000 (scaled by a 256 MiB device copy in the same run) from
# the rocprofv3 --pmc runs summarised in profiles/pmc_single_job_L501_m4_N1024.md
TRAFFIC_BYTES = {"offspring": int((1952.2 * 2.000 + 5865.2) * 1024),
                 "stem": int((816.7 * 2.000 + 82.6) * 1024)}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stem_cells(L):
    n = 3 * L
    return (n - 5) * (n - 4) // 2


def child_bytes(L, m):
    """Algorithmic HBM bytes of one child in the offspring launch: the parent's
    genome read, the survivor copy and the child written (one byte per codon
    each), and ~96 bytes of objective and stem records."""
    return 3 * m * L + 96


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
    ap.add_argument("--batched-gens", type=int, default=20,
                    help="generations timed in the batched-vs-per-instance comparison (0 skips it)")
    ap.add_argument("--batched-jobs", type=int, default=64)
    ap.add_argument("--served", type=float, default=0.0, metavar="SECONDS",
                    help="instead of the single-job line, run the served form: open-loop Poisson arrivals for "
                         "this many seconds (the served workload's mix) and print its own JSON line")
    ap.add_argument("--rate", type=float, default=20.0, help="arrival rate of --served, jobs/s")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB three times on the GPU (a plain copy "
                         "kernel of known size, to calibrate hardware byte counters); not part of the measurement")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus must be 1 "
                                   "(jobs would shard across GPUs by job, with no communication)"}))
        sys.exit(2)

    if args.served > 0:
        print(json.dumps(bench_served(importlib.import_module(PKG), args)), flush=True)
        return
    cdp, aa, tab = workload(args)
    L = len(aa)
    default_workload = (args.residues, args.m, args.N, args.H, args.Pm, args.seed) == (500, 4, 1024, 5, 0.1, 2)
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
    cell_rate = cells / (prof[1] / 1e3)
    log(f"per-kernel profile: {prof}")
    dominant = KERNELS[int(np.argmax(prof))]
    stem_line = {"bound": "valu", "achieved": round(cell_rate / 1e12, 4), "peak": round(PEAK_CELLS / 1e12, 2),
                 "unit": "T stem cells/s (peak: 32 cells per int32 lane-op)", "frac": round(cell_rate / PEAK_CELLS, 5),
                 "traffic": TRAFFIC_BYTES["stem"] if default_workload else None, "kernel": "k_stem_list",
                 "cells_per_launch": int(cells)}
    if dominant == "stem":
        roofline = stem_line
    else:  # the offspring launch (k_design), the dominant kernel at the default workload
        nbytes = args.N * child_bytes(L, args.m)
        rate = nbytes / (prof[0] / 1e3)
        roofline = {"bound": "hbm", "achieved": round(rate / 1e9, 2), "peak": HBM_BYTES_S / 1e9, "unit": "GB/s",
                    "frac": round(rate / HBM_BYTES_S, 5),
                    "traffic": TRAFFIC_BYTES["offspring"] if default_workload else None,
                    "kernel": "k_design<offspring>", "bytes_per_launch": int(nbytes),
                    "dominant_in_profile": dominant}
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
        "roofline": roofline,
        "stem_scan": stem_line,
    }
    if args.batched_gens > 0:
        out["batched"] = bench_batched(cdp, args)
        log(f"batched: {out['batched']}")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(aa, tab, args)
        log(f"cpu baseline: {out['cpu_baseline']['value']:.1f} designs/s")
    job.close()
    print(json.dumps(out), flush=True)
    if args.copy_mib > 0:
        copy_calibration(args.copy_mib)


def bench_batched(cdp, args):
    """Mixed-size batched workload (64 jobs: U{50..2000} aa, m in {2,4,8},
    N in {256,512,1024}, H=4, master seed 4), all resident at once: the same
    jobs timed advanced one at a time (each job's own graph on its own
    stream, one after another) and advanced together (one graph per
    generation over all of them).  Results are identical either way
    (tests/test_gpu_batch.py); only the time differs."""
    W = importlib.import_module(PKG + ".workloads")
    specs = W.batched_mix(args.batched_jobs, 4)
    jobs = [cdp.Job(*sp.build()).init() for sp in specs]
    K = args.batched_gens
    for j in jobs:
        j.step(2)
    for j in jobs:
        j.sync()
    log(f"batched mix: {len(jobs)} jobs created and initialised")
    per_gen = []  # per-instance: latency of one generation of one job
    t_alone = 0.0
    for j in jobs:
        ms = [j.step_timed(1) for _ in range(K)]
        per_gen += ms
        t_alone += sum(ms)
    b = cdp.Batch(jobs)
    b.step(2)
    b.sync()
    bat = [b.step_timed(1) for _ in range(K)]
    t_batch = sum(bat)
    b.close()
    for j in jobs:
        j.close()
    designs = sum(sp.N for sp in specs) * K
    return {"workload": f"{len(specs)} jobs, U{{50..2000}} aa, m in {{2,4,8}}, N in {{256,512,1024}}, H=4, seed 4; "
                        f"{K} timed generations",
            "designs": designs,
            "per_instance_designs_s": round(designs / (t_alone / 1e3), 1),
            "batched_designs_s": round(designs / (t_batch / 1e3), 1),
            "speedup": round(t_alone / t_batch, 3),
            "per_instance_generation_ms": {"p50": round(float(np.percentile(per_gen, 50)), 4),
                                           "p99": round(float(np.percentile(per_gen, 99)), 4)},
            "batched_generation_ms": {"p50": round(float(np.percentile(bat, 50)), 4),
                                      "p99": round(float(np.percentile(bat, 99)), 4)}}


def bench_served(cdp, args):
    """Served form: Poisson arrivals at --rate jobs/s for --served seconds
    (lognormal protein length, median 350 aa, clipped to 30..3000; m in
    {2,4,8}; N in {256..2048}; 300..1000 generations; H=4; master seed 5)
    submitted to one server with a 128 GiB device budget.  Reports designs/s
    over the whole run and p50/p99 time to front (submission to result on
    the host), beside the same jobs run sequentially one at a time through
    the per-instance path (timed on every 10th job and scaled, stated)."""
    import threading
    W = importlib.import_module(PKG + ".workloads")
    arrivals = W.served_arrivals(args.rate, args.served, 5)
    built = [a.spec.build() for a in arrivals]
    log(f"served: {len(arrivals)} arrivals over {args.served:.0f} s")
    srv = cdp.Server(128 << 30, 512)
    tickets = [None] * len(arrivals)
    t0 = time.perf_counter()

    def submitter():
        for i, a in enumerate(arrivals):
            dt = a.t - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            tickets[i] = srv.submit(*built[i], a.spec.generations)

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    log("served: all submitted")
    ttf, queued = [], []
    for i, t in enumerate(tickets):
        while srv.wait(t, 30.0) not in (cdp.SERVER_DONE, cdp.SERVER_FAILED):
            log(f"served: waiting for job {i} of {len(tickets)}; {srv.stats()}")
        r = srv.result(t)
        ttf.append(r["times"]["time_to_front"])
        queued.append(r["times"]["queued"])
    wall = time.perf_counter() - t0
    designs = sum(a.spec.N * a.spec.generations for a in arrivals)
    srv.close()
    log(f"served: done in {wall:.1f} s")
    # the same jobs one at a time through the per-instance path (a sample)
    seq_ms, seq_designs = 0.0, 0
    for i in range(0, len(arrivals), 10):
        j = cdp.Job(*built[i]).init()
        seq_ms += j.step_timed(arrivals[i].spec.generations)
        seq_designs += arrivals[i].spec.N * arrivals[i].spec.generations
        j.close()
    seq_rate = seq_designs / (seq_ms / 1e3)
    log(f"served: sequential sample done")
    return {"metric": "served form: designs/s and p50/p99 time-to-front under open-loop Poisson arrivals",
            "value": round(designs / wall, 1), "unit": "designs/s", "n_gpus": 1, "higher_is_better": True,
            "data": "synthetic",
            "config": {"workload": f"Poisson arrivals at {args.rate:g} jobs/s for {args.served:g} s: lognormal "
                                   f"protein length (median 350 aa, sigma 0.8, clipped 30..3000), m in {{2,4,8}}, "
                                   f"N in {{256..2048}}, 300..1000 generations, H=4, master seed 5; 128 GiB budget",
                       "jobs": len(arrivals), "designs": designs},
            "wall_s": round(wall, 3),
            "time_to_front_s": {"p50": round(float(np.percentile(ttf, 50)), 3),
                                "p99": round(float(np.percentile(ttf, 99)), 3),
                                "max": round(float(np.max(ttf)), 3)},
            "queued_s": {"p50": round(float(np.percentile(queued, 50)), 3),
                         "p99": round(float(np.percentile(queued, 99)), 3)},
            "sequential_per_instance": {"designs_s": round(seq_rate, 1),
                                        "sample": f"every 10th job ({len(range(0, len(arrivals), 10))} jobs) "
                                                  f"run alone, GPU time by HIP events",
                                        "estimated_wall_s": round(designs / seq_rate, 1)},
            "speedup_vs_sequential": round((designs / wall) / seq_rate, 3)}


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
