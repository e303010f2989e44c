"""Batched form: uneven jobs advanced together by one launch sequence give
exactly the pools each job gives alone (and the oracle gives)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_jobs(cdp, specs):
    T = cdp.tables
    out = []
    for i, (res, m, N, H, seed) in enumerate(specs):
        aa = T.synth_protein(res, seed)
        tab = T.synth_tables(seed)
        out.append((aa, tab, m, N, H, 0.1, seed))
    return out


SPECS = [(50, 2, 32, 2, 11), (120, 4, 64, 3, 12), (37, 8, 16, 1, 13), (300, 2, 128, 4, 14), (75, 3, 48, 2, 15)]


def test_batch_equals_alone(cdp, orc):
    print("batch", flush=True)
    specs = make_jobs(cdp, SPECS)
    G = 7
    alone = []
    for s in specs:
        j = cdp.Job(*s).init()
        j.step(G)
        alone.append(j.read())
        j.close()
    jobs = [cdp.Job(*s).init() for s in specs]
    b = cdp.Batch(jobs)
    b.step(3)
    b.step(G - 3)
    b.sync()
    for j, a, s in zip(jobs, alone, specs):
        r = j.read()
        for k in ("genome", "obj", "idx", "pql", "sorted", "front_of", "extremes", "nadir"):
            assert np.array_equal(r[k], a[k]), k
    # and the oracle agrees with one of them
    aa, tab, m, N, H, Pm, seed = specs[1]
    o = orc.Job(aa, tab, m, N, H, Pm, seed)
    o.init()
    o.step(G)
    ro = o.read()
    for k in ("genome", "obj", "sorted", "front_of"):
        assert np.array_equal(jobs[1].read()[k], ro[k]), k
    b.close()
    for j in jobs:
        j.close()


def test_batch_two_lds_classes(cdp):
    """A batch whose jobs fall in both k_design launch classes (the 1300-residue
    job with m=8 needs about 26 KiB of LDS, more than six workgroups per CU
    allow; the others fit six) and are reordered longest first gives every job
    exactly its pools alone."""
    print("batch two LDS classes", flush=True)
    specs = make_jobs(cdp, [(60, 2, 32, 2, 31), (1300, 8, 16, 2, 32), (900, 4, 24, 2, 33), (200, 3, 40, 3, 34)])
    G = 3
    alone = []
    for s in specs:
        j = cdp.Job(*s).init()
        j.step(G)
        alone.append(j.read())
        j.close()
    jobs = [cdp.Job(*s).init() for s in specs]
    b = cdp.Batch(jobs)
    b.step(G)
    b.sync()
    for i, (j, a) in enumerate(zip(jobs, alone)):
        r = j.read()
        for k in ("genome", "obj", "idx", "pql", "sorted", "front_of"):
            assert np.array_equal(r[k], a[k]), (i, k)
    b.close()
    for j in jobs:
        j.close()


def test_batch_rejects_uninitialised(cdp):
    specs = make_jobs(cdp, SPECS[:2])
    j0 = cdp.Job(*specs[0]).init()
    j1 = cdp.Job(*specs[1])
    with pytest.raises(cdp.CdError):
        cdp.Batch([j0, j1])
    with pytest.raises(cdp.CdError):
        cdp.Batch([j0, j0])


def test_cds_scanned_counts(cdp):
    """The count of CDS whose objectives were recomputed since initialisation:
    per generation at most every CDS of the children (one to m per child,
    fewer when an operator finds nothing to change); the same whether a job
    advances alone or in a batch."""
    print("cds scanned", flush=True)
    specs = make_jobs(cdp, SPECS)
    G = 5
    alone = []
    for s in specs:
        m, N = s[2], s[3]
        j = cdp.Job(*s).init()
        j.sync()
        n0 = j.cds_scanned()
        assert n0 == 0
        prev = n0
        for _ in range(G):
            j.step(1)
            j.sync()
            n = j.cds_scanned()
            assert 0 <= n - prev <= N * m
            prev = n
        assert prev > n0  # the children did change CDS
        alone.append(prev)
        j.close()
    jobs = [cdp.Job(*s).init() for s in specs]
    b = cdp.Batch(jobs)
    b.step(G)
    b.sync()
    assert [j.cds_scanned() for j in jobs] == alone
    b.close()
    for j in jobs:
        j.close()


KEYS = ("genome", "obj", "idx", "pql", "sorted", "front_of", "extremes", "nadir")


def test_block_graphs_equal_single_generations(cdp, orc):
    """Steps of eight generations or more launch a graph that holds eight
    generations back to back.  Each replay gives exactly the pools of
    single-generation graph launches and of eager launches (profile), read
    back between replays, and the oracle agrees after all of them."""
    print("block graphs: jobs", flush=True)
    specs = make_jobs(cdp, [SPECS[1], SPECS[3]])
    total = 0
    for si, s in enumerate(specs):
        blk, one, eager = (cdp.Job(*s).init() for _ in range(3))
        total = 0
        for G in (8, 8, 11, 16, 3):
            blk.step(G)
            for _ in range(G):
                one.step(1)
            eager.profile(G)
            total += G
            rb, r1, re = blk.read(), one.read(), eager.read()
            for k in KEYS:
                assert np.array_equal(rb[k], r1[k]), (si, total, k)
                assert np.array_equal(rb[k], re[k]), (si, total, k)
        if si == 0:
            o = orc.Job(*s)
            o.init()
            o.step(total)
            ro = o.read()
            for k in ("genome", "obj", "sorted", "front_of"):
                assert np.array_equal(blk.read()[k], ro[k]), k
        for j in (blk, one, eager):
            j.close()
    print("block graphs: batch", flush=True)
    specs = make_jobs(cdp, SPECS)
    alone = []
    for s in specs:
        j = cdp.Job(*s).init()
        for _ in range(19):
            j.step(1)
        alone.append(j.read())
        j.close()
    jobs = [cdp.Job(*s).init() for s in specs]
    b = cdp.Batch(jobs)
    b.step(19)
    b.sync()
    for i, (j, a) in enumerate(zip(jobs, alone)):
        r = j.read()
        for k in KEYS:
            assert np.array_equal(r[k], a[k]), (i, k)
    b.close()
    for j in jobs:
        j.close()
