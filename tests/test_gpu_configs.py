"""Each benchmark configuration at its own workload on the GPU.

* The single large job (1500 residues, m = 8, N = 4096, H = 5, seed 3): two
  generations of the loop at that shape.  Children are checked against the
  oracle's mutation of their parents and its evaluation (a sample), the
  whole pool's fronts against the oracle's sort, and the selection against
  the oracle's selection on the same pool and extremes.
* The batched mix (64 jobs, U{50..2000} residues, m in {2,4,8}, N in
  {256,512,1024}, H = 4): three generations batched equal three generations
  of every job alone, and the oracle agrees on the two smallest jobs.
* The served mix (Poisson arrivals, lognormal lengths up to 3000 residues,
  N up to 2048): a slice of the arrival process with short generation
  budgets, every result equal to the same job run alone.
* Dynamic LDS limits are process-wide: a smaller job created after a larger
  one must not lower them under the larger one's launches.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("genome", "obj", "idx", "pql", "sorted", "front_of")


def _say(*a):
    print(*a, flush=True)


def _equal(a, b, keys, what):
    for k in keys:
        assert np.array_equal(a[k], b[k]), (what, k, np.argwhere(np.asarray(a[k]) != np.asarray(b[k]))[:5])


def test_large_job_two_generations(cdp, orc):
    _say("large job: 1500 residues, m=8, N=4096, H=5, seed 3")
    T = cdp.tables
    aa, tab = T.synth_protein(1500, 3), T.synth_tables(3)
    m, N, H, Pm, seed = 8, 4096, 5, 0.1, 3
    job = cdp.Job(aa, tab, m, N, H, Pm, seed).init()
    r0 = job.read()
    job.step(1)
    r1 = job.read()
    job.step(1)
    r2 = job.read()
    job.close()
    _say("large job: generations run")
    rng = np.random.default_rng(3)
    for gen, (prev, cur) in enumerate(((r0, r1), (r1, r2))):
        srt = prev["sorted"]
        # survivors: copies of the parents the previous selection kept
        assert np.array_equal(cur["genome"][:N], prev["genome"][srt])
        assert np.array_equal(cur["obj"][:N], prev["obj"][srt])
        # a sample of children: the oracle's mutation of the parent, then its evaluation
        kids = np.sort(rng.choice(N, size=16, replace=False))
        for c in kids:
            p = srt[c]
            want = orc.mutate(aa, tab, Pm, seed, gen, int(c), prev["genome"][p], prev["obj"][p], prev["idx"][p],
                              prev["pql"][p])
            assert np.array_equal(cur["genome"][N + c], want), (gen, int(c))
        obj, idx, pql = orc.evaluate(aa, tab, cur["genome"][N + kids])
        assert np.array_equal(obj, cur["obj"][N + kids]) and np.array_equal(idx, cur["idx"][N + kids])
        assert np.array_equal(pql, cur["pql"][N + kids])
        _say("large job: generation", gen, "children match")
    # fronts of the whole pool (8192 designs) and the selection of generation 1
    fo, nf = orc.nds(r2["obj"], N)
    assert np.array_equal(fo, r2["front_of"])
    sel = orc.select(r2["obj"], N, H, 1, seed, r1["extremes"])
    assert np.array_equal(sel["sorted"], r2["sorted"]) and np.array_equal(sel["front_of"], r2["front_of"])
    for k in ("extremes", "ideal", "worst", "nadir"):
        assert np.array_equal(sel[k], r2[k]), k
    _say("large job: fronts and selection match;", nf, "fronts")


def test_batched_mix_equals_alone(cdp, orc):
    W = cdp.workloads
    specs = W.batched_mix(64, 4)
    G = 3
    _say("batched mix: 64 jobs alone")
    alone = []
    for s in specs:
        j = cdp.Job(*s.build()).init()
        j.step(G)
        alone.append(j.read())
        j.close()
    _say("batched mix: 64 jobs batched")
    jobs = [cdp.Job(*s.build()).init() for s in specs]
    b = cdp.Batch(jobs)
    b.step(G)
    b.sync()
    for i, (j, a) in enumerate(zip(jobs, alone)):
        _equal(j.read(), a, KEYS, ("job", i))
    b.close()
    for j in jobs:
        j.close()
    _say("batched mix: equal; oracle on the two smallest jobs")
    small = sorted(range(len(specs)), key=lambda i: specs[i].residues * specs[i].m * specs[i].N)[:2]
    for i in small:
        o = orc.Job(*specs[i].build())
        o.init()
        o.step(G)
        _equal(o.read(), alone[i], KEYS, ("oracle", i))


def served_slice(W):
    arr = W.served_arrivals(20.0, 600.0, 5)
    pick = list(range(20))
    pick.append(max(range(len(arr)), key=lambda i: (arr[i].spec.residues, arr[i].spec.N)))
    pick.append(next(i for i in range(len(arr)) if arr[i].spec.N == 2048))
    return [arr[i].spec for i in pick]


def test_served_slice_equals_alone(cdp):
    W = cdp.workloads
    specs = served_slice(W)
    assert max(s.residues for s in specs) == 3000 and max(s.N for s in specs) == 2048
    gens = [2 + (i % 3) for i in range(len(specs))]
    _say("served slice:", len(specs), "jobs alone")
    want = []
    for s, G in zip(specs, gens):
        j = cdp.Job(*s.build()).init()
        j.step(G)
        want.append(j.read())
        j.close()
    _say("served slice: served")
    srv = cdp.Server(64 << 30)
    tickets = [srv.submit(*s.build(), G) for s, G in zip(specs, gens)]
    for i, t in enumerate(tickets):
        assert srv.wait(t, 120.0) == cdp.SERVER_DONE, specs[i]
        _equal(srv.result(t), want[i], KEYS, ("served", i))
    srv.close()
    _say("served slice: equal")


def test_lds_limits_survive_a_smaller_job(cdp, orc):
    """k_niche needs 94 KiB for P = 4096 and 51 KiB for P = 2048; creating the
    smaller job after the larger one must leave the larger one runnable and
    its results unchanged."""
    _say("lds limits: large job then smaller job")
    T = cdp.tables
    big = (T.synth_protein(80, 41), T.synth_tables(41), 2, 2048, 3, 0.1, 41)
    small = (T.synth_protein(60, 42), T.synth_tables(42), 2, 1024, 3, 0.1, 42)
    j = cdp.Job(*big).init()
    j.step(2)
    want = j.read()
    j.close()
    jb = cdp.Job(*big).init()
    js = cdp.Job(*small).init()
    jb.step(2)
    _equal(jb.read(), want, KEYS, "large after small")
    js.step(1)
    js.sync()
    jb.close()
    js.close()


def test_batch_holds_its_jobs(cdp):
    T = cdp.tables
    args = (T.synth_protein(40, 43), T.synth_tables(43), 2, 16, 2, 0.1, 43)
    j = cdp.Job(*args).init()
    b = cdp.Batch([j])
    with pytest.raises(cdp.CdError, match="CD_E_STATE"):
        j.step(1)
    assert cdp.lib().cd_job_destroy(j.h) == -4
    b.step(2)
    assert j.generation == 2  # waits for the batch's stream
    b.close()
    j.step(1)
    assert j.generation == 3
    j.close()
