"""Equal hypervolume: over 10 seeds, the final pool's hypervolume (main.cu's
normalisation, reference point 1^6, exact WFG) is the same for the engine
run alone, the engine in a batch, and the CPU oracle.  The pools are
bit-identical, so the hypervolumes agree exactly, well inside the 1 %
tolerance the comparison asks for."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEEDS = range(31, 41)
RES, M, N, H, G = 60, 2, 32, 2, 12


def test_final_hypervolume_equal_over_seeds(cdp, orc):
    print("hv seeds", flush=True)
    T = cdp.tables
    specs = [(T.synth_protein(RES, s), T.synth_tables(s), M, N, H, 0.1, s) for s in SEEDS]
    hv_alone, hv_batch, hv_oracle = [], [], []
    for sp in specs:
        j = cdp.Job(*sp).init()
        j.step(G)
        hv_alone.append(cdp.quality(j.read()["obj"])[0])
        j.close()
        o = orc.Job(*sp)
        o.init()
        o.step(G)
        hv_oracle.append(cdp.quality(o.read()["obj"])[0])
    jobs = [cdp.Job(*sp).init() for sp in specs]
    b = cdp.Batch(jobs)
    b.step(G)
    b.sync()
    hv_batch = [cdp.quality(j.read()["obj"])[0] for j in jobs]
    b.close()
    for j in jobs:
        j.close()
    assert hv_alone == hv_oracle
    assert hv_batch == hv_alone
    assert all(h > 0.0 for h in hv_alone)
    mean = float(np.mean(hv_alone))
    assert abs(float(np.mean(hv_batch)) - mean) <= 0.01 * mean
