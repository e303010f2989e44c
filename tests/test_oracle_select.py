"""Pins of the oracle's selection (include/sorting.cuh).

Fronts are checked against repeated extraction of the eps-non-dominated set,
the hyperplane solve against numpy.linalg.solve, extremes against a direct
ASF argmin, association against FP64 perpendicular distances, and niching
against the NSGA-III balance property and the level-0 rule (a reference line
with no member in the earlier fronts takes its nearest candidate first).  The
level-0 check is also run against an oracle built with a planted mistake.
"""
import numpy as np
import pytest

EPS = np.float32(1e-6)


def dom(a, b):  # sorting.cuh:94-130 restated on numpy float32
    d = np.abs(a - b) > EPS
    return bool(d.any() and not (a[d] > b[d]).any())


def brute_fronts(obj, N):
    P = len(obj)
    left = set(range(P))
    fronts = []
    ranked = 0
    while ranked < N and left:
        f = sorted(j for j in left if not any(dom(obj[i], obj[j]) for i in left if i != j))
        if not f:
            f = sorted(left)
        fronts.append(f)
        left -= set(f)
        ranked += len(f)
    return fronts


def quantised_pool(rng, P):
    obj = rng.normal(size=(P, 6)).astype(np.float32)
    obj[:, 2] = np.round(obj[:, 2] * 3) / 3  # discrete objectives produce ties
    obj[:, 5] = np.round(obj[:, 5] * 2) / 2
    dup = rng.integers(0, P, size=P // 8)
    obj[dup] = obj[rng.integers(0, P, size=len(dup))]
    near = rng.integers(0, P, size=P // 8)
    obj[near, 0] += np.float32(5e-7)  # below the eps threshold
    return obj


def test_nds_matches_brute_force(orc):
    rng = np.random.default_rng(1)
    for trial in range(25):
        N = int(rng.integers(3, 40))
        obj = quantised_pool(rng, 2 * N)
        if trial % 3 == 0:
            obj[:, 1:] = obj[:, :1] * 0 + rng.normal(size=(2 * N, 5)).astype(np.float32) * 0.01
            obj[:, 0] = np.arange(2 * N, dtype=np.float32)  # long chains of fronts
        fo, nf = orc.nds(obj, N)
        want = brute_fronts(obj, N)
        assert nf == len(want)
        for f, members in enumerate(want):
            assert sorted(np.nonzero(fo == f)[0].tolist()) == members
        assert (fo >= 0).sum() >= N


def asf(x, ideal, i):
    t = x - ideal
    t = np.where(t.astype(np.float64) < 0.001, np.float32(0), t).astype(np.float32)
    w = np.full(6, np.float32(1e-6), np.float32)
    w[i] = 1
    return (t / w).max()


def test_selection_invariants(orc):
    rng = np.random.default_rng(2)
    for trial in range(20):
        N = int(rng.integers(6, 60))
        H = int(rng.integers(1, 4))
        obj = quantised_pool(rng, 2 * N)
        ext0 = rng.normal(size=(6, 6)).astype(np.float32)
        r = orc.select(obj, N, H, trial, 99, ext0)
        fo = r["front_of"]
        srt = r["sorted"]
        assert len(set(srt.tolist())) == N and (np.diff(srt) > 0).all()
        last = fo.max()
        assert all(fo[i] >= 0 for i in srt)
        assert set(np.nonzero((fo >= 0) & (fo < last))[0]) <= set(srt.tolist())
        assert np.array_equal(r["ideal"], obj.min(axis=0)) and np.array_equal(r["worst"], obj.max(axis=0))
        # extremes: sequential argmin of ASF over F0 (ascending) then the previous extremes
        F0 = np.nonzero(fo == 0)[0]
        ext = ext0.copy()
        for i in range(6):
            cand = [obj[j] for j in F0] + [ext[e] for e in range(6)]
            vals = [asf(c, r["ideal"], i) for c in cand]
            ext[i] = cand[int(np.argmin(vals))]
        assert np.array_equal(ext, r["extremes"])
        # nadir: hyperplane intercepts when accepted, else max over F0 (or worst)
        if r["hyp_exception"]:
            mx = obj[F0].max(axis=0)
            want = np.where(mx - r["ideal"] > EPS, mx, r["worst"])
            assert np.array_equal(r["nadir"], want)
        else:
            x = np.linalg.solve(ext.astype(np.float64), np.ones(6))
            assert np.allclose(r["nadir"], r["ideal"] + 1 / x, rtol=1e-3)


def test_hyperplane_solve(orc):
    """Extremes a_i e_i with ideal 0: the hyperplane test passes and the
    intercepts (sorting.cuh:911-972) are the a_i."""
    rng = np.random.default_rng(3)
    N = 12
    a = rng.uniform(0.5, 2, size=6).astype(np.float32)
    obj = np.concatenate([np.diag(a), rng.uniform(0.2, 0.4, size=(2 * N - 6, 6)).astype(np.float32)])
    r = orc.select(obj, N, 2, 0, 1, np.full((6, 6), 10, np.float32))
    assert np.array_equal(r["extremes"], np.diag(a))
    assert r["hyp_exception"] == 0
    assert np.allclose(r["nadir"], a, rtol=1e-6)
    # shifting the pool moves the ideal off the origin; the reference's test
    # (result = x.(E - ideal) == 1) then fails and MNDF takes over
    r2 = orc.select(obj + np.float32(0.25), N, 2, 0, 1, np.full((6, 6), 10, np.float32))
    assert r2["hyp_exception"] == 1


def test_niche_balance(orc):
    rng = np.random.default_rng(4)
    seen_niching = 0
    for trial in range(30):
        N = int(rng.integers(10, 80))
        H = int(rng.integers(1, 4))
        obj = rng.normal(size=(2 * N, 6)).astype(np.float32)
        obj[:, 3] = -obj[:, 0] - obj[:, 1]  # many mutually non-dominated designs
        r = orc.select(obj, N, H, trial, 7, np.zeros((6, 6), np.float32))
        if r["n_cut"]:
            continue
        seen_niching += 1
        fo, srt = r["front_of"], set(r["sorted"].tolist())
        last = fo.max()
        ideal, nadir = r["ideal"].astype(np.float64), r["nadir"].astype(np.float64)
        den = nadir - ideal
        refs = orc.dasdennis(6, H).astype(np.float64)

        def assoc(i):
            f = np.where(den == 0, 0, (obj[i] - ideal) / np.where(den == 0, 1, den))
            k = (refs @ f) / (refs * refs).sum(axis=1)
            d = np.sqrt(((k[:, None] * refs - f) ** 2).sum(axis=1))
            return int(np.argmin(d)), d
        counts = np.zeros(len(refs), int)
        cand = {}
        near_tie = False
        for i in np.nonzero(fo >= 0)[0]:
            a, d = assoc(i)
            if len(d) > 1 and np.sort(d)[1] - d[a] < 1e-5:  # float32 vs float64 may disagree
                near_tie = True
                break
            if fo[i] < last:
                counts[a] += 1
            else:
                cand.setdefault(a, []).append(i)
        if near_tie:
            continue
        final = counts.copy()
        for a, lst in cand.items():
            final[a] += sum(1 for i in lst if i in srt)
        picked = [a for a, lst in cand.items() if any(i in srt for i in lst)]
        leftover = [a for a, lst in cand.items() if any(i not in srt for i in lst)]
        if picked and leftover:
            assert max(final[a] for a in picked) <= min(final[a] for a in leftover) + 1
    assert seen_niching > 5


def level0_check(orc, trials=40):
    """referenceBasedSorting's level 0 (sorting.cuh:1191-1245): a reference
    point with no associated design in the fronts before the last takes, as
    its first pick from the last front, the candidate nearest to its line.
    So whenever any candidate of such a point survives, its nearest one
    does.  Association and distances in float64; cases where float32 could
    order two lines or two candidates differently are skipped."""
    rng = np.random.default_rng(11)
    checked = 0
    for trial in range(trials):
        N = int(rng.integers(20, 120))
        H = int(rng.integers(2, 5))
        obj = rng.normal(size=(2 * N, 6)).astype(np.float32)
        obj[:, 3] = -obj[:, 0] - obj[:, 1]  # a large first front: most lines start at count 0
        r = orc.select(obj, N, H, trial, 5, np.zeros((6, 6), np.float32))
        if r["n_cut"]:
            continue
        fo, srt = r["front_of"], set(r["sorted"].tolist())
        last = fo.max()
        ideal, nadir = r["ideal"].astype(np.float64), r["nadir"].astype(np.float64)
        den = nadir - ideal
        refs = orc.dasdennis(6, H).astype(np.float64)
        ranked = np.nonzero(fo >= 0)[0]
        f = np.where(den == 0, 0, (obj[ranked] - ideal) / np.where(den == 0, 1, den))
        k = (f @ refs.T) / (refs * refs).sum(axis=1)[None, :]
        d = np.sqrt(((k[:, :, None] * refs[None] - f[:, None, :]) ** 2).sum(axis=2))
        order = np.sort(d, axis=1)
        if len(refs) > 1 and (order[:, 1] - order[:, 0] < 1e-5).any():
            continue
        line = dict(zip(ranked.tolist(), np.argmin(d, axis=1).tolist()))
        dist = dict(zip(ranked.tolist(), order[:, 0].tolist()))
        count = np.zeros(len(refs), int)
        for i, a in line.items():
            if fo[i] < last:
                count[a] += 1
        for a in range(len(refs)):
            cand = [i for i, b in line.items() if b == a and fo[i] == last]
            if count[a] or len(cand) < 2 or not any(i in srt for i in cand):
                continue
            ds = sorted(dist[i] for i in cand)
            if ds[1] - ds[0] < 1e-5 * max(1.0, ds[0]):
                continue
            nearest = min(cand, key=lambda i: dist[i])
            assert nearest in srt, (trial, a, nearest, sorted(cand))
            checked += 1
    assert checked >= 50, checked


def test_niche_level0_takes_the_nearest(orc):
    level0_check(orc)


def test_niche_level0_check_catches_planted_mistake(orc, tmp_path):
    path = str(tmp_path / "libcd_oracle_level0.so")
    assert orc.build_variant(path, [("void select_pool(", "if (level == 0)", "if (level < 0)")]).wait() == 0
    with orc.use_library(path):
        with pytest.raises(AssertionError):
            level0_check(orc)
