"""Benchmark of the codon-design engine on one MI355X.

Headline (the JSON line's `value`): the batched mixed-size workload -- 64
independent jobs, protein lengths U{50..2000} residues, m in {2,4,8} CDS,
populations N in {256,512,1024}, Das-Dennis H = 4, 500 generations each,
per-job synthetic tables, master seed 4 -- all resident in HBM and advanced
together, one hipGraph-launched generation (a "tick") per step.  `value` is
designs/s (sum of N over the jobs x K ticks / time of the K timed ticks, HIP
events on the batch's stream, after W warmup ticks).  Beside it:

* per_instance: the same jobs advanced one at a time, each with its own
  graph on its own stream, K generations back to back per job (the plain
  per-instance GPU loop), and the speedup of the batch over it; both sides
  run the launch shapes the autotuner picked for them (cd_job_tune on each
  job alone, cd_batch_tune on the batch: ordinary generations of the jobs,
  timed per kernel), reported with the batch's untuned tick;
* c3_full: both forms over the workload's whole 500 generations from
  freshly initialised jobs, with designs/s, time-to-front p50/p99 and the
  final pools' hypervolume on both sides (the pools are bit-identical);
* c1: the single job (500 residues, m = 4, N = 1024, H = 5, seed 2) over its
  full 1000 generations; c2: the single large job (1500 residues, m = 8,
  N = 4096, H = 5, seed 3) over its full 2000 generations;
* roofline: the dominant kernel of the batched tick (its per-phase profile,
  HIP events on the batch's stream) against its HBM roofline, with the HBM
  traffic and the VALU / LDS counters of the rocprofv3 runs summarised in
  profiles/;
* cpu_baseline: the CPU oracle compiled on this host with -march=native, on
  a bounded sample of the batched workload; cpu_c0: the CPU-runnable single
  job (100 residues, m = 2, N = 128, H = 4, 100 generations, seed 1) run to
  completion on all cores and on one core.

`--served SECONDS` prints the served form's line instead (Poisson arrivals).
"""
import argparse
import concurrent.futures as cf
import importlib
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
PKG = "protein_encoding_v2_nsga-iii_amd"

METRIC = "designs/s + p50/p99 latency, batched vs per-instance on MI355X at equal HV (±1%)"
HBM_BYTES_S = 8.0e12            # MI355X HBM3E peak
VALU_WAVE_INSTR_S = 256 * 2.4e9  # 256 CUs x 4 SIMDs x one wave64 VALU instruction per 4 cycles, 2.4 GHz
KERNELS = ("offspring", "stem", "dominance", "fronts", "associate", "niche")
SECTIONS = ("headline", "c3full", "c1", "c2", "cpu")
DEFAULT_BATCH_SHAPES = dict(design_split=24 * 1024, design_threads=512, stem_waves=7, dom_rows=128,
                            fronts_threads=1024, assoc_designs=4, niche_threads=1024)

# Hardware counters of the offspring phase of one batched tick at the default
# workload (both k_design<2> launch classes), from the rocprofv3 --pmc runs
# summarised in profiles/pmc_batched_mix_64jobs.md: HBM bytes (FETCH_SIZE x
# its copy-kernel factor + WRITE_SIZE x its factor), VALU wave-instructions,
# LDS instructions and LDS bank-conflict cycles.
TICK_COUNTERS = {"hbm_bytes": 736134804, "valu_wave_instr": 244725985, "lds_instr": 24882951,
                 "lds_conflict_cycles": 6257523}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def child_bytes(L, m):
    """Algorithmic HBM bytes of one child in the offspring kernel: the parent's
    genome read, the survivor copy and the child written (one byte per codon
    each), and ~96 bytes of objective and stem records."""
    return 3 * m * L + 96


def pct(x, q):
    return round(float(np.percentile(np.asarray(x, np.float64), q)), 4)


def hv_of(cdp, objs, threads=16):
    """Final-pool hypervolume of each pool (main.cu's normalisation, reference
    point 1^6, exact WFG), computed on host threads."""
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(lambda o: cdp.quality(o)[0], objs))


def build_jobs(cdp, specs):
    return [cdp.Job(*s.build()).init() for s in specs]


def bench_headline(cdp, args):
    W = cdp.workloads
    specs = W.batched_mix(64, 4)
    designs_per_tick = sum(s.N for s in specs)
    jobs = build_jobs(cdp, specs)
    for j in jobs:
        j.sync()
    log(f"batched mix: {len(jobs)} jobs created and initialised")
    # the plain per-instance loop: each job alone (its launch shapes tuned on
    # its own generations), K generations back to back
    for j in jobs:
        j.tune(1)
        j.step(args.warmup)
    per_ms = [j.step_timed(args.steps) for j in jobs]
    t_alone = sum(per_ms)
    log(f"per-instance: {args.steps} generations of each job alone, {t_alone:.1f} ms in all")
    b = cdp.Batch(jobs)
    untuned = b.shapes(set=DEFAULT_BATCH_SHAPES)
    b.step(2)
    b.sync()
    t_untuned = b.step_timed(5) / 5
    shapes, tune_gens = b.tune(3)  # three ticks a candidate: the candidates differ by about the ticks' spread
    log(f"batch shapes tuned over {tune_gens} generations: {shapes} (untuned tick {t_untuned:.3f} ms)")
    b.step(args.warmup)
    b.sync()
    t_batch = b.step_timed(args.steps)
    log(f"batched: {args.steps} ticks in {t_batch:.2f} ms")
    ticks = [b.step_timed(1) for _ in range(min(args.steps, 50))]
    prof = [float(x) for x in b.profile(10)]
    log(f"batched tick profile (ms): {dict(zip(KERNELS, [round(x, 4) for x in prof]))}")
    b.close()
    for j in jobs:
        j.close()
    nbytes = sum(s.N * child_bytes(s.residues + 1, s.m) for s in specs)
    dominant = KERNELS[int(np.argmax(prof))]
    roof = None
    if dominant == "offspring":
        rate = nbytes / (prof[0] / 1e3)
        roof = {"bound": "hbm", "achieved": round(rate / 1e9, 2), "peak": HBM_BYTES_S / 1e9, "unit": "GB/s",
                "frac": round(rate / HBM_BYTES_S, 5),
                "traffic": TICK_COUNTERS["hbm_bytes"] if TICK_COUNTERS else None,
                "kernel": "k_design<offspring>: both launch classes of one tick",
                "bytes_per_launch": int(nbytes), "ms_per_launch": round(prof[0], 4)}
    else:
        roof = {"bound": "see kernel", "kernel": dominant, "ms_per_launch": round(prof[KERNELS.index(dominant)], 4),
                "achieved": None, "peak": None, "unit": None, "frac": None, "traffic": None}
    counters = None
    if TICK_COUNTERS:
        c = TICK_COUNTERS
        counters = {"kernel": "k_design<offspring> (one tick's two launches)",
                    "valu_issue_frac": round(c["valu_wave_instr"] / (prof[0] / 1e3) / VALU_WAVE_INSTR_S, 4),
                    "valu_peak_wave_instr_s": VALU_WAVE_INSTR_S,
                    "lds_bank_conflict_cycles_per_lds_instr": round(c["lds_conflict_cycles"] / c["lds_instr"], 4),
                    "source": "profiles/pmc_batched_mix_64jobs.md"}
    return {
        "value": round(designs_per_tick * args.steps / (t_batch / 1e3), 1),
        "ms_per_step": round(t_batch / args.steps, 4),
        "config": {"workload": "batched mix: 64 jobs resident, U{50..2000} residues, m in {2,4,8}, "
                               "N in {256,512,1024}, H=4, Pm=0.1, master seed 4; one step = one hipGraph-launched "
                               "generation of all 64 jobs",
                   "jobs": len(specs), "designs_per_step": designs_per_tick,
                   "parallelism": "single GPU, all jobs batched per launch sequence"},
        "latency_ms_per_generation": {"batched_p50": pct(ticks, 50), "batched_p99": pct(ticks, 99),
                                      "per_instance_p50": pct(np.asarray(per_ms) / args.steps, 50),
                                      "per_instance_p99": pct(np.asarray(per_ms) / args.steps, 99)},
        "per_instance": {"designs_s": round(designs_per_tick * args.steps / (t_alone / 1e3), 1),
                         "ms_for_all_jobs": round(t_alone, 3),
                         "how": "each job alone on its own stream, K generations back to back (HIP events), summed"},
        "speedup_vs_per_instance": round(t_alone / t_batch, 3),
        "autotuner": {"shapes": shapes, "tuning_generations": tune_gens, "untuned_shapes": untuned,
                      "untuned_ms_per_step": round(t_untuned, 4)},
        "kernel_ms": dict(zip(KERNELS, [round(x, 4) for x in prof])),
        "roofline": roof,
        "counters": counters,
    }


def bench_c3_full(cdp, args):
    W = cdp.workloads
    specs = W.batched_mix(64, 4)
    G = specs[0].generations
    designs = sum(s.N for s in specs) * G
    # per-instance: one job after another, each from initialisation to its last generation
    alone_ms, alone_obj = [], []
    for s in specs:
        j = cdp.Job(*s.build()).init()
        j.sync()
        alone_ms.append(j.step_timed(G))
        alone_obj.append(j.read()["obj"])
        j.close()
    log(f"c3 full: per-instance {sum(alone_ms):.0f} ms")
    jobs = build_jobs(cdp, specs)
    b = cdp.Batch(jobs)
    b.sync()
    t_batch = b.step_timed(G)
    b.close()
    batch_obj = [j.read()["obj"] for j in jobs]
    for j in jobs:
        j.close()
    log(f"c3 full: batched {t_batch:.0f} ms")
    same = all(np.array_equal(a, c) for a, c in zip(alone_obj, batch_obj))
    hv_a = hv_of(cdp, alone_obj)
    hv_b = hv_of(cdp, batch_obj)
    ttf_alone = np.cumsum(alone_ms) / 1e3  # sequential: a job's front is ready when it and those before it ran
    return {"workload": "the batched mix over its whole 500 generations, from initialisation (init not timed)",
            "generations": G, "designs": designs,
            "batched_designs_s": round(designs / (t_batch / 1e3), 1),
            "per_instance_designs_s": round(designs / (sum(alone_ms) / 1e3), 1),
            "speedup": round(sum(alone_ms) / t_batch, 3),
            "time_to_front_s": {"batched_p50": round(t_batch / 1e3, 4), "batched_p99": round(t_batch / 1e3, 4),
                                "per_instance_sequential_p50": pct(ttf_alone, 50),
                                "per_instance_sequential_p99": pct(ttf_alone, 99)},
            "hypervolume_mean": {"batched": float(np.mean(hv_b)), "per_instance": float(np.mean(hv_a)),
                                 "rel_diff": float(abs(np.mean(hv_b) - np.mean(hv_a)) / np.mean(hv_a))},
            "pools_bit_identical": bool(same)}


def bench_single(cdp, residues, m, N, H, seed, G, label):
    T = cdp.tables
    aa, tab = T.synth_protein(residues, seed), T.synth_tables(seed)
    t = cdp.Job(aa, tab, m, N, H, 0.1, seed).init()  # a copy to tune on; the shapes are cached
    # eight generations a candidate: a lone generation's phases differ by a few microseconds between
    # shapes, about the spread of two generations' timings
    shapes, _ = t.tune(8)
    t.close()
    log(f"{label}: shapes tuned: {shapes}")
    j = cdp.Job(aa, tab, m, N, H, 0.1, seed).init()
    assert j.shapes() == shapes
    j.sync()
    ms = j.step_timed(G)
    lat = [j.step_timed(1) for _ in range(50)]
    prof = [float(x) for x in j.profile(10)]
    hv = cdp.quality(j.read()["obj"])[0]
    j.close()
    log(f"{label}: {G} generations in {ms:.1f} ms")
    return {"workload": f"single job: {residues} residues (L={residues + 1} with the stop), m={m}, N={N} "
                        f"(pool {2 * N}), H={H}, Pm=0.1, seed {seed}; {G} generations after initialisation",
            "designs_s": round(N * G / (ms / 1e3), 1), "ms_per_generation": round(ms / G, 4),
            "latency_ms_per_generation": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
            "kernel_ms": dict(zip(KERNELS, [round(x, 4) for x in prof])), "shapes": shapes,
            "final_hypervolume": hv}


def cpu_runs(cdp, args):
    import oracle
    threads = min(os.cpu_count() or 1, 16)
    path = oracle.build_native(os.path.join(tempfile.mkdtemp(prefix="cd_oracle_"), "libcd_oracle_native.so"))
    build = "g++ -O3 -march=native -fopenmp on this host"
    if path is None:
        path, build = oracle.LIB, "the shipped -O3 build (-march=native failed here)"
    T = cdp.tables
    out = {}
    with oracle.use_library(path):
        # the CPU-runnable single job to completion, all cores and one core
        aa, tab = T.synth_protein(100, 1), T.synth_tables(1)
        c0 = {}
        for label, n in (("all_cores", threads), ("one_core", 1)):
            oracle.set_threads(n)
            job = oracle.Job(aa, tab, 2, 128, 4, 0.1, 1)
            t0 = time.perf_counter()
            job.init()
            t1 = time.perf_counter()
            job.step(100)
            t2 = time.perf_counter()
            c0[label] = {"designs_s": round(128 * 100 / (t2 - t1), 1), "generations_s": round(t2 - t1, 3),
                         "init_s": round(t1 - t0, 3), "cores": n}
        out["cpu_c0"] = {"workload": "single job: 100 residues, m=2, N=128, H=4, Pm=0.1, 100 generations, seed 1 "
                                     "(generation loop timed as main.cu:707-731, init apart)", "build": build, **c0}
        log(f"cpu c0: {c0}")
        # the headline workload, bounded: the 8 smallest jobs of the mix, 2 generations each
        oracle.set_threads(threads)
        specs = sorted(cdp.workloads.batched_mix(64, 4), key=lambda s: (s.residues + 1) * s.m * s.N)[:8]
        gens, dt = 2, 0.0
        for s in specs:
            job = oracle.Job(*s.build())
            job.init()
            t0 = time.perf_counter()
            job.step(gens)
            dt += time.perf_counter() - t0
        designs = sum(s.N for s in specs) * gens
        out["cpu_baseline"] = {"value": round(designs / dt, 1), "unit": "designs/s", "cores": threads, "kind": "port",
                               "sample": f"the 8 smallest jobs of the batched mix ({designs} designs), 2 generations "
                                         f"each after initialisation, {dt:.1f} s; oracle built {build}, OpenMP over "
                                         f"children"}
        log(f"cpu baseline: {out['cpu_baseline']['value']} designs/s")
    oracle.set_threads(threads)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed batched ticks")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--sections", default=",".join(SECTIONS),
                    help="comma-separated parts to run: " + ",".join(SECTIONS)
                    + " (the headline runs unless --no-headline)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--served", type=float, default=0.0, metavar="SECONDS",
                    help="instead of the batched line, run the served form: open-loop Poisson arrivals for "
                         "this many seconds and print its own JSON line")
    ap.add_argument("--rate", type=float, default=20.0, help="arrival rate of --served, jobs/s")
    ap.add_argument("--no-headline", action="store_true",
                    help="skip the batched headline and print only the sections asked for (to profile one of them "
                         "alone); the line then carries no value")
    ap.add_argument("--single-gens", type=int, default=0,
                    help="generations of the c1 and c2 sections (default: their full runs of 1000 and 2000); "
                         "for shorter profiling runs")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB three times on the GPU (a plain copy "
                         "kernel of known size, to calibrate hardware byte counters); not part of the measurement")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus must be 1 "
                                   "(jobs would shard across GPUs by job, with no communication)"}))
        sys.exit(2)
    cdp = importlib.import_module(PKG)
    if args.served > 0:
        print(json.dumps(bench_served(cdp, args)), flush=True)
        return
    sections = set(x.strip() for x in args.sections.split(",") if x.strip())
    if args.no_headline:
        out = {"metric": METRIC, "value": None, "unit": "designs/s", "n_gpus": 1, "note": "headline skipped"}
    else:
        out = headline_line(bench_headline(cdp, args), args)
    if "c3full" in sections:
        out["c3_full"] = bench_c3_full(cdp, args)
    if "c1" in sections:
        out["c1"] = bench_single(cdp, 500, 4, 1024, 5, 2, args.single_gens or 1000, "c1")
    if "c2" in sections:
        out["c2"] = bench_single(cdp, 1500, 8, 4096, 5, 3, args.single_gens or 2000, "c2")
    if "cpu" in sections and not args.no_cpu:
        out.update(cpu_runs(cdp, args))
    print(json.dumps(out), flush=True)
    if args.copy_mib > 0:
        copy_calibration(args.copy_mib)


def headline_line(head, args):
    """The JSON line of the batched headline."""
    return {
        "metric": METRIC,
        "value": head["value"],
        "unit": "designs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 codons, int32/int64 integer objectives, fp32 CAI/CPB",
        "data": "synthetic",
        "config": head["config"],
        "latency_ms_per_generation": head["latency_ms_per_generation"],
        "per_instance": head["per_instance"],
        "speedup_vs_per_instance": head["speedup_vs_per_instance"],
        "kernel_ms": head["kernel_ms"],
        "roofline": head["roofline"],
        "counters": head["counters"],
    }


def bench_served(cdp, args):
    """Served form: Poisson arrivals at --rate jobs/s for --served seconds
    (lognormal protein length, median 350 aa, clipped to 30..3000; m in
    {2,4,8}; N in {256..2048}; 300..1000 generations; H=4; master seed 5)
    submitted to one server with a 128 GiB device budget.  Reports designs/s
    over the whole run and p50/p99 time to front (submission to result on
    the host), beside the same jobs run sequentially one at a time through
    the per-instance path (timed on every 10th job and scaled, stated)."""
    import threading
    W = cdp.workloads
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
    st = srv.stats()
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
    log("served: sequential sample done")
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
            "ticks": st["generations"],
            "tick_gpu_s": round(st["tick_gpu_s"], 3),
            "gpu_busy_frac": round(st["tick_gpu_s"] / wall, 4),
            "resident_set_changes": st["rebuilds"], "rebuild_s": round(st["rebuild_s"], 3),
            "worker": {"busy_s": round(st["worker_busy_s"], 3), "launch_s": round(st["launch_s"], 3),
                       "wait_s": round(st["wait_s"], 3)},
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
