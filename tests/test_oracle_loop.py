"""Pins of the oracle's generation loop (src/main.cu:121-240, 700-732).

Each child's mutation type is recomputed from its Philox stream and the child
is checked against what that operator (include/mutation.cuh) may change.
The operators' choice rules and Bernoulli rates are then restated
independently in numpy over recorded parent and child pools: which codons
are eligible, what a change must achieve, the rate at which eligible codons
change, and the uniformity of random picks.  Every such check is also run
against an oracle built with a deliberately planted mistake and must fail
there (the mistakes are planted on the CPU, in a private build).
"""
import os

import numpy as np
import pytest
import scipy.stats

TAG_TYPE = 0x54595045


def mutation_type(orc, seed, gen, child):
    x = int(orc.philox([gen, child, 0, TAG_TYPE], [seed & 0xFFFFFFFF, seed >> 32])[0])
    return (x * 7) >> 32


@pytest.mark.parametrize("tables_kind", ["reference", "synthetic"])
def test_loop_operator_invariants(cdp, orc, tables_kind):
    T = cdp.tables
    tables = T.reference_tables() if tables_kind == "reference" else T.synth_tables(4)
    aa = T.synth_protein(60, 7)
    L, m, N, seed = len(aa), 3, 48, 12345
    job = orc.Job(aa, tables, m, N, 2, 0.1, seed)
    job.init()
    r0 = job.read()
    g0 = r0["genome"]
    assert (g0[0] == np.array([T.SYN_NUM[a] - 1 for a in aa])[None, :]).all()  # HIGHEST_CAI_GEN
    assert r0["obj"][0, 0] == -1.0 or tables_kind == "synthetic"
    seen = set()
    for gen in range(6):
        prev = job.read()
        job.step(1)
        cur = job.read()
        G, Gp = cur["genome"], prev["genome"]
        ns = np.array([T.SYN_NUM[a] for a in aa])
        assert ((G >= 0) & (G < ns[None, None, :])).all()
        for c in range(N):
            par = prev["sorted"][c]
            assert np.array_equal(G[c], Gp[par]) and np.array_equal(cur["obj"][c], prev["obj"][par])
            child, parent = G[N + c], Gp[par]
            pidx, ppql = prev["idx"][par], prev["pql"][par]
            t = mutation_type(orc, seed, gen, c)
            seen.add(t)
            diff = np.argwhere(child != parent)
            cds = set(diff[:, 0].tolist())
            if t == 1:
                assert cds <= {pidx[0]} and all(child[i, k] > parent[i, k] for i, k in diff)
            elif t == 2:
                assert cds <= {pidx[2]} and all(k % 2 == gen % 2 and k < L - 1 for _, k in diff)
                for i, k in diff:
                    gl = lambda s, kk: T.SYN_START[aa[kk]] + int(s[kk])
                    def score(s):
                        v = np.float32(0)
                        if k > 0:
                            v = np.float32(v + tables.cps[gl(s, k - 1) * 61 + gl(s, k)])
                        if gl(s, k + 1) < 61:
                            v = np.float32(v + tables.cps[gl(s, k) * 61 + gl(s, k + 1)])
                        return v
                    assert score(child[i]) > score(parent[i])
            elif t == 3:
                assert cds <= {pidx[4]} and all(k % 2 == gen % 2 for _, k in diff)
            elif t == 4:
                assert cds <= {pidx[6]}
            elif t == 5:
                assert cds <= {pidx[8]}
                for i, k in diff:
                    g3 = lambda s: T.decode(aa[k:k + 1], tables, s[k:k + 1])[2] in "GC"
                    assert g3(child[i]) != g3(parent[i])
                    assert g3(child[i]) == (pidx[9] == 1)
            elif t == 6:
                assert len(diff) <= 1 and cds <= {pidx[10]}
                if len(diff):
                    l_idx, r_idx = ppql[0] // 3, (ppql[0] + ppql[2] - 1) // 3
                    assert diff[0][1] == (l_idx + r_idx) // 2
        # stored records equal a fresh evaluation of the stored genome
        obj, idx, pql = orc.evaluate(aa, tables, G)
        assert np.array_equal(obj, cur["obj"]) and np.array_equal(idx, cur["idx"]) and np.array_equal(pql, cur["pql"])
        srt = cur["sorted"]
        assert len(set(srt.tolist())) == N and (cur["front_of"][srt] >= 0).all()
    assert seen == set(range(7))


def test_loop_deterministic(cdp, orc, ref_tables):
    T = cdp.tables
    aa = T.synth_protein(30, 2)
    runs = []
    for seed in (5, 5, 6):
        j = orc.Job(aa, ref_tables, 2, 16, 2, 0.2, seed)
        j.init()
        j.step(5)
        runs.append(j.read())
    assert np.array_equal(runs[0]["genome"], runs[1]["genome"])
    assert not np.array_equal(runs[0]["genome"], runs[2]["genome"])


def test_invalid_jobs_rejected(cdp, orc, ref_tables):
    T = cdp.tables
    aa = T.synth_protein(30, 2)
    for kw in (dict(m=1), dict(N=0), dict(Pm=1.5), dict(Pm=-0.1)):
        args = dict(m=2, N=8, H=2, Pm=0.1, seed=1)
        args.update(kw)
        with pytest.raises(ValueError):
            orc.Job(aa, ref_tables, args["m"], args["N"], args["H"], args["Pm"], args["seed"])
    bad = aa.copy()
    bad[3] = 20  # a stop inside the protein
    with pytest.raises(ValueError):
        orc.Job(bad, ref_tables, 2, 8, 2, 0.1, 1)


# ---------------------------------------------------------------------------
# Operator rules and rates over recorded pools
# ---------------------------------------------------------------------------
PM, SEED, M, NPOP, GENS = 0.2, 777, 4, 128, 8
Z_MAX = 4.5  # binomial z bound of the rate checks
P_MIN = 1e-6  # chi-square p-value bound of the uniformity checks

# deliberate mistakes: (function, text, replacement) in oracle/cd_oracle.cpp
PLANTS = {
    "random_keeps_equal_draw": [("void mut_random(", "if (nw >= cur) ++nw;", "if (nw > cur) ++nw;")],
    "cai_always_top": [("void mut_cai(", "cur = (int8_t)(cur + 1 + (int)uniform_int(pick, (uint32_t)(ns - 1 - cur)));",
                        "cur = (int8_t)(ns - 1);")],
    "cbp_rate_pm": [("void mut_cbp(", "c.thr_2pm", "c.thr_pm")],
    "hsc_rate_pm": [("void mut_hsc(", "c.thr_2pm", "c.thr_pm")],
    "hsc_accepts_ties": [("void mut_hsc(", "count(base + j) > curv", "count(base + j) >= curv")],
    "hd_first_improving": [("void mut_hd(", "if (nm > cur_mhd && nm > best_mhd) {",
                            "if (nm > cur_mhd && best_mhd == -1) {")],
    "gc3_twice_deviation": [("void mut_gc3(", "bern_threshold(dev)", "bern_threshold(2.0f * dev)")],
    "sl_rate_2pm": [("void mut_sl(", "bern(r[0], c.thr_pm)", "bern(r[0], c.thr_2pm)")],
    "niche_level0_by_priority": [("void select_pool(", "if (level == 0)", "if (level < 0)")],
}


@pytest.fixture(scope="session")
def planted(tmp_path_factory, orc):
    d = tmp_path_factory.mktemp("planted")
    procs = {k: orc.build_variant(str(d / f"libcd_oracle_{k}.so"), v) for k, v in PLANTS.items()}
    for k, p in procs.items():
        assert p.wait() == 0, k
    return {k: str(d / f"libcd_oracle_{k}.so") for k in PLANTS}


_records = {}


def recorded_loop(cdp, orc):
    """Parent and child pools of GENS generations of one job (150 residues,
    m = 4, N = 128, Pm = 0.2) on the oracle build in use."""
    key = orc.lib()._name
    if key not in _records:
        T = cdp.tables
        aa, tab = T.synth_protein(150, 21), T.synth_tables(4)
        job = orc.Job(aa, tab, M, NPOP, 3, PM, SEED)
        job.init()
        prev, recs = job.read(), []
        for gen in range(GENS):
            job.step(1)
            cur = job.read()
            recs.append((gen, prev, cur))
            prev = cur
        _records[key] = (aa, tab, recs)
    return _records[key]


class Ctx:
    """Codon facts of one job computed from the codon table as text."""

    def __init__(self, cdp, aa, tab):
        T = cdp.tables
        self.aa, self.tab, self.L = aa, tab, len(aa)
        self.base = np.array([T.SYN_START[a] for a in aa])
        self.ns = np.array([T.SYN_NUM[a] for a in aa])
        self.cod = [tab.codons[3 * g:3 * g + 3] for g in range(64)]
        self.stops = {self.cod[g] for g in (61, 62, 63)}
        self.cps = np.asarray(tab.cps, np.float32)

    def g(self, cds, k):  # global codon index of a synonym row
        return int(self.base[k] + cds[k])

    def hidden(self, a, b):  # stop triplets read across the junction of codons a|b
        x, y = self.cod[a], self.cod[b]
        return int(x[1:] + y[0] in self.stops) + int(x[2] + y[:2] in self.stops)

    def ham(self, a, b):
        return sum(p != q for p, q in zip(self.cod[a], self.cod[b]))


def children(cdp, orc, kind):
    """(gen, ctx, parent genome, child genome, parent record) of every child whose operator is `kind`."""
    aa, tab, recs = recorded_loop(cdp, orc)
    ctx = Ctx(cdp, aa, tab)
    for gen, prev, cur in recs:
        for c in range(NPOP):
            if mutation_type(orc, SEED, gen, c) != kind:
                continue
            p = prev["sorted"][c]
            rec = dict(obj=prev["obj"][p], idx=prev["idx"][p], pql=prev["pql"][p])
            yield gen, ctx, prev["genome"][p], cur["genome"][NPOP + c], rec


def assert_rate(changed, eligible, p, what):
    n = len(eligible) if np.ndim(p) else eligible
    if np.ndim(p):  # per-trial probabilities
        p = np.asarray(p, np.float64)
        mean, var = p.sum(), (p * (1 - p)).sum()
    else:
        mean, var = p * n, n * p * (1 - p)
    assert var > 0, what
    z = (changed - mean) / np.sqrt(var)
    assert abs(z) < Z_MAX, (what, changed, mean, z)


def assert_uniform(groups, what):
    """groups: {k: list of picks in 0..k-1}; every group with data must look uniform."""
    tested = 0
    for k, picks in groups.items():
        if k < 2 or len(picks) < 20 * k:
            continue
        obs = np.bincount(picks, minlength=k)
        assert scipy.stats.chisquare(obs).pvalue > P_MIN, (what, k, obs.tolist())
        tested += 1
    assert tested, (what, "no group large enough")


def check_random(cdp, orc):
    """mutationRandom (mutation.cuh:16-63): every codon of every CDS with more
    than one synonym changes with probability Pm, to one of the others
    uniformly."""
    n_el = n_ch = 0
    groups = {}
    for gen, ctx, par, ch, rec in children(cdp, orc, 0):
        multi = ctx.ns > 1
        diff = ch != par
        assert not (diff & ~multi[None, :]).any()
        n_el += int(multi.sum()) * par.shape[0]
        n_ch += int(diff.sum())
        for c, k in np.argwhere(diff):
            ns, cur, new = int(ctx.ns[k]), int(par[c, k]), int(ch[c, k])
            groups.setdefault(ns - 1, []).append(new if new < cur else new - 1)
    assert_rate(n_ch, n_el, PM, "random rate")
    assert_uniform(groups, "random pick")


def check_cai(cdp, orc):
    """mutationCAI, SELECT_UPPER_RANDOM (mutation.cuh:65-142): in the worst-CAI
    CDS, a codon below the last (best) synonym changes with probability Pm to
    a uniformly chosen higher one."""
    n_el = n_ch = 0
    groups = {}
    for gen, ctx, par, ch, rec in children(cdp, orc, 1):
        c = rec["idx"][0]
        diff = ch != par
        assert not np.delete(diff, c, axis=0).any()
        up = par[c] < ctx.ns - 1
        assert not (diff[c] & ~up).any() and (ch[c][diff[c]] > par[c][diff[c]]).all()
        n_el += int(up.sum())
        n_ch += int(diff[c].sum())
        for k in np.nonzero(diff[c])[0]:
            groups.setdefault(int(ctx.ns[k] - 1 - par[c, k]), []).append(int(ch[c, k] - par[c, k] - 1))
    assert_rate(n_ch, n_el, PM, "cai rate")
    assert_uniform(groups, "cai pick")


def local_candidates(ctx, cds, k, score, strict_better):
    cur = int(cds[k])
    vals = [score(int(ctx.base[k]) + j) for j in range(ctx.ns[k])]
    return [j for j in range(ctx.ns[k]) if j != cur and strict_better(vals[j], vals[cur])]


def check_parity_operator(cdp, orc, kind):
    """mutationCBP (mutation.cuh:144-262) and mutationHSC (264-388),
    SELECT_UPPER_RANDOM: in the worst CDS, codons of the generation's parity
    (CBP: not the last two; HSC: every one up to the stop) that have a
    synonym with a strictly better local score (CBP: the pair scores with
    both neighbours, no left at codon 0, no right into the stop; HSC: hidden
    stops across both junctions) change with probability 2 Pm, uniformly
    among those synonyms."""
    n_el = n_ch = 0
    groups = {}
    for gen, ctx, par, ch, rec in children(cdp, orc, kind):
        c = rec["idx"][2 if kind == 2 else 4]
        diff = ch != par
        assert not np.delete(diff, c, axis=0).any()
        s = par[c]
        last = ctx.L - 1 if kind == 2 else ctx.L
        for k in range(ctx.L):
            ok = k % 2 == gen % 2 and k < last and ctx.ns[k] > 1
            if not ok:
                assert not diff[c, k]
                continue
            left = ctx.g(s, k - 1) if k > 0 else None
            right = ctx.g(s, k + 1) if k + 1 < ctx.L else None
            if kind == 2:
                def score(g):
                    v = np.float32(0)
                    if left is not None:
                        v = np.float32(v + ctx.cps[left * 61 + g])
                    if right < 61:
                        v = np.float32(v + ctx.cps[g * 61 + right])
                    return v
            else:
                def score(g):
                    return (ctx.hidden(left, g) if left is not None else 0) + \
                        (ctx.hidden(g, right) if right is not None else 0)
            cand = local_candidates(ctx, s, k, score, lambda a, b: a > b)
            if not cand:
                assert not diff[c, k], ("changed without a strictly better synonym", kind, gen, k)
                continue
            n_el += 1
            if diff[c, k]:
                n_ch += 1
                assert int(ch[c, k]) in cand, (kind, gen, k)
                groups.setdefault(len(cand), []).append(cand.index(int(ch[c, k])))
    assert_rate(n_ch, n_el, 2 * PM, ("parity operator rate", kind))
    if kind == 2:
        assert_uniform(groups, "cbp pick")


def check_hd(cdp, orc):
    """mutationHD (mutation.cuh:390-534): codons of the first CDS of the worst
    pair change with probability Pm when a synonym raises the codon's
    smallest Hamming distance to the other CDS (then to the largest such
    distance), or keeps it and raises the distance to the pair's second CDS
    (then to the largest such distance)."""
    n_el = n_ch = 0
    for gen, ctx, par, ch, rec in children(cdp, orc, 4):
        c1, c2 = int(rec["idx"][6]), int(rec["idx"][7])
        diff = ch != par
        assert not np.delete(diff, c1, axis=0).any()
        others = [o for o in range(par.shape[0]) if o != c1]
        for k in range(ctx.L):
            if ctx.ns[k] == 1:
                assert not diff[c1, k]
                continue
            gs = [int(ctx.base[k]) + j for j in range(ctx.ns[k])]
            mhd = [min(ctx.ham(g, ctx.g(par[o], k)) for o in others) for g in gs]
            pair = [ctx.ham(g, ctx.g(par[c2], k)) for g in gs]
            cur = int(par[c1, k])
            alt = [j for j in range(len(gs)) if j != cur]
            best_m = max(mhd[j] for j in alt)
            if best_m > mhd[cur]:
                allowed = {j for j in alt if mhd[j] == best_m}
            else:
                tied = [j for j in alt if mhd[j] == mhd[cur] and pair[j] > pair[cur]]
                allowed = {j for j in tied if pair[j] == max(pair[t] for t in tied)} if tied else set()
            if not allowed:
                assert not diff[c1, k]
                continue
            n_el += 1
            if diff[c1, k]:
                n_ch += 1
                assert int(ch[c1, k]) in allowed, (gen, k, mhd, pair, cur, int(ch[c1, k]))
    assert_rate(n_ch, n_el, PM, "hd rate")


def check_gc3(cdp, orc):
    """mutationGC3 (mutation.cuh:536-653): in the worst-GC3 CDS, a codon with
    a synonym whose third base moves G/C content the needed way changes to
    such a synonym with probability equal to the parent's GC3 deviation."""
    probs, n_ch = [], 0
    for gen, ctx, par, ch, rec in children(cdp, orc, 5):
        c, up = int(rec["idx"][8]), int(rec["idx"][9]) == 1
        dev = float(rec["obj"][4])
        diff = ch != par
        assert not np.delete(diff, c, axis=0).any()
        for k in range(ctx.L):
            gc = [ctx.cod[int(ctx.base[k]) + j][2] in "GC" for j in range(ctx.ns[k])]
            cur = int(par[c, k])
            ok = [j for j in range(ctx.ns[k]) if j != cur and (gc[j] > gc[cur] if up else gc[j] < gc[cur])]
            if not ok:
                assert not diff[c, k]
                continue
            probs.append(dev)
            if diff[c, k]:
                n_ch += 1
                assert int(ch[c, k]) in ok
    assert_rate(n_ch, probs, probs, "gc3 rate")


def check_sl(cdp, orc):
    """mutationSL (mutation.cuh:775-969): the first trial, at the middle codon
    of the parent's stem in its CDS, changes that codon with probability Pm
    and ends the operator (its accept test compares against an integer
    quotient, mutation.cuh:843)."""
    n_el = n_ch = 0
    for gen, ctx, par, ch, rec in children(cdp, orc, 6):
        c = int(rec["idx"][10])
        p, l = int(rec["pql"][0]), int(rec["pql"][2])
        diff = ch != par
        assert not np.delete(diff, c, axis=0).any() and diff[c].sum() <= 1
        mid = (p // 3 + (p + l - 1) // 3) // 2
        if l == 0 or ctx.ns[mid] == 1:
            continue
        n_el += 1
        if diff[c].any():
            assert np.nonzero(diff[c])[0].tolist() == [mid]
            n_ch += 1
    assert_rate(n_ch, n_el, PM, "sl rate")


CHECKS = {
    "random": (check_random, "random_keeps_equal_draw"),
    "cai": (check_cai, "cai_always_top"),
    "cbp": (lambda cdp, orc: check_parity_operator(cdp, orc, 2), "cbp_rate_pm"),
    "hsc_rate": (lambda cdp, orc: check_parity_operator(cdp, orc, 3), "hsc_rate_pm"),
    "hsc_choice": (lambda cdp, orc: check_parity_operator(cdp, orc, 3), "hsc_accepts_ties"),
    "hd": (check_hd, "hd_first_improving"),
    "gc3": (check_gc3, "gc3_twice_deviation"),
    "sl": (check_sl, "sl_rate_2pm"),
}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_operator_rule_holds(cdp, orc, name):
    CHECKS[name][0](cdp, orc)


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_operator_rule_catches_planted_mistake(cdp, orc, planted, name):
    check, plant = CHECKS[name]
    with orc.use_library(planted[plant]):
        with pytest.raises(AssertionError):
            check(cdp, orc)
