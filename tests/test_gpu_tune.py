"""Autotuner: the search runs ordinary generations in every candidate launch
shape, so a tuned job or batch ends exactly where it would have ended
untuned (and where each job ends alone); the chosen shapes are cached for
later work of the same kind."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("genome", "obj", "idx", "pql", "sorted", "front_of")
# (residues, m, N, H, seed): both k_design LDS classes, N up to 512
SPECS = [(50, 2, 32, 2, 51), (1300, 8, 16, 2, 52), (300, 4, 512, 4, 53), (120, 3, 64, 3, 54)]


def build(cdp, spec):
    res, m, N, H, seed = spec
    T = cdp.tables
    return (T.synth_protein(res, seed), T.synth_tables(seed), m, N, H, 0.1, seed)


def alone(cdp, spec, G):
    j = cdp.Job(*build(cdp, spec)).init()
    j.shapes(set=dict(design_split=0, design_threads=256, stem_waves=6, dom_rows=64, fronts_threads=512,
                      assoc_designs=4, niche_threads=1024))
    j.step(G)
    r = j.read()
    j.close()
    return r


def test_batch_tune_keeps_results(cdp):
    print("batch tune", flush=True)
    jobs = [cdp.Job(*build(cdp, s)).init() for s in SPECS]
    b = cdp.Batch(jobs)
    chosen, run = b.tune(1)
    assert run >= 19  # every candidate of every parameter ran a generation
    assert b.shapes() == chosen
    b.step(2)
    b.sync()
    for j, s in zip(jobs, SPECS):
        want = alone(cdp, s, run + 2)
        got = j.read()
        for k in KEYS:
            assert np.array_equal(got[k], want[k]), (s, k)
    b.close()
    # a later batch of the same kind starts from the cached shapes
    jobs2 = [cdp.Job(*build(cdp, s)).init() for s in SPECS]
    b2 = cdp.Batch(jobs2)
    assert b2.shapes() == chosen
    b2.close()
    for j in jobs + jobs2:
        j.close()


def test_job_tune_keeps_results(cdp):
    print("job tune", flush=True)
    spec = SPECS[2]
    j = cdp.Job(*build(cdp, spec)).init()
    chosen, run = j.tune(1)
    assert run >= 15
    j.step(2)
    want = alone(cdp, spec, run + 2)
    got = j.read()
    for k in KEYS:
        assert np.array_equal(got[k], want[k]), k
    j.close()


def test_invalid_shapes_rejected(cdp):
    j = cdp.Job(*build(cdp, SPECS[0])).init()
    good = j.shapes()
    for k, v in (("stem_waves", 9), ("dom_rows", 96), ("fronts_threads", 128), ("assoc_designs", 2),
                 ("niche_threads", 2048), ("design_split", -1), ("design_threads", 384)):
        bad = dict(good, **{k: v})
        with pytest.raises(cdp.CdError, match="CD_E_INVALID"):
            j.shapes(set=bad)
    assert j.shapes() == good
    j.close()


@pytest.mark.parametrize("assoc", [32, 64])
def test_association_shapes_keep_results(cdp, assoc):
    """Eight lanes per design in k_associate (32 or 64 designs a workgroup)
    choose the same reference lines as a wave per design."""
    print("association shape", assoc, flush=True)
    for spec in (SPECS[2], SPECS[3]):
        want = alone(cdp, spec, 4)
        j = cdp.Job(*build(cdp, spec)).init()
        j.shapes(set=dict(design_split=0, design_threads=256, stem_waves=6, dom_rows=64, fronts_threads=512,
                          assoc_designs=assoc, niche_threads=1024))
        j.step(4)
        got = j.read()
        for k in KEYS:
            assert np.array_equal(got[k], want[k]), (spec, k)
        j.close()
