"""The engine's quality indicators (host code in libcd_engine.so; no GPU):
exact hypervolume against inclusion-exclusion and Monte Carlo, the
reference's normalisation (main.cu:789-835) and MinEuclid."""
import itertools

import numpy as np
import pytest


def hv_inclusion_exclusion(pts, ref):
    total = 0.0
    n = len(pts)
    for r in range(1, n + 1):
        for sub in itertools.combinations(range(n), r):
            corner = np.max(pts[list(sub)], axis=0)
            total += (-1) ** (r + 1) * np.prod(np.clip(ref - corner, 0, None))
    return total


@pytest.mark.parametrize("d", [2, 3, 4, 6])
def test_hv_exact_small(cdp, d):
    rng = np.random.default_rng(d)
    for trial in range(15):
        n = int(rng.integers(1, 10))
        pts = rng.random((n, d))
        if trial % 3 == 0:
            pts = np.round(pts * 4) / 4  # ties and duplicates
        want = hv_inclusion_exclusion(pts, np.ones(d))
        assert abs(cdp.hypervolume(pts) - want) < 1e-12


def test_hv_closed_forms(cdp):
    assert cdp.hypervolume(np.zeros((1, 6))) == 1.0
    assert cdp.hypervolume(np.full((3, 6), 0.5)) == pytest.approx(0.5 ** 6)
    assert cdp.hypervolume(np.array([[1.0, 0.0]])) == 0.0  # on the reference boundary
    # the unit simplex corners: 6 boxes of volume 1*... pairwise overlapping
    e = np.eye(6) * 0 + (1 - np.eye(6)) * 0.0
    e[np.arange(6), np.arange(6)] = 0.5
    assert cdp.hypervolume(e) == pytest.approx(hv_inclusion_exclusion(e, np.ones(6)))


def test_hv_monte_carlo(cdp):
    rng = np.random.default_rng(9)
    pts = rng.random((60, 6)) ** 0.5
    s = rng.random((400000, 6))
    dominated = np.zeros(len(s), bool)
    for p in pts:
        dominated |= np.all(s >= p, axis=1)
    assert abs(cdp.hypervolume(pts) - dominated.mean()) < 4 * np.sqrt(0.25 / len(s))


def test_normalize_and_min_distance(cdp):
    raw = np.array([[-1.0, -0.2, -0.4, -0.5, 0.0, 0.0],     # the ideal point -> origin
                    [0.0, 0.15, 0.0, 0.0, 0.6, 1.0],        # the nadir -> clipped to 0.999999
                    [-0.5, 0.0, -0.2, -0.25, 0.3, 0.5]], np.float32)
    z = cdp.normalize(raw)
    assert np.array_equal(z[0], np.zeros(6, np.float32))
    assert np.allclose(z[1], 0.999999)
    assert np.allclose(z[2], [0.5, 0.2 / 0.35, 0.5, 0.5, 0.5, 0.5], atol=1e-6)
    hv, md, sp = cdp.quality(raw)
    assert md == 0.0 and sp == pytest.approx(np.sqrt(6) * 0.999999, rel=1e-6)
    assert hv == pytest.approx(1.0)
