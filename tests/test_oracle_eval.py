"""Pins of the oracle's objective evaluators (include/utils.cuh:271-888).

Every check restates the objective differently from the oracle's code:
hidden stops as off-frame stop triplets in the nucleotide string, Hamming
distance on strings, GC3 on every third base, CAI/CPB in FP64 closed form and
the stem by brute force, plus hand-built hairpins and the B4TWR7 fixture.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")
STOPS = {"UAA", "UAG", "UGA"}
COMP = {"A": "U", "U": "A", "G": "C", "C": "G"}


def brute_objectives(cdp, aa, tables, design):
    T = cdp.tables
    m, L = design.shape
    n = 3 * L
    seqs = [T.decode(aa, tables, design[c]) for c in range(m)]
    g = [[T.SYN_START[a] + int(s) for a, s in zip(aa, design[c])] for c in range(m)]
    cai = [np.exp(np.mean([np.log(np.float64(tables.weights[x])) for x in gc])) for gc in g]
    cpb = [np.mean([np.float64(tables.cps[gc[k] * 61 + gc[k + 1]]) for k in range(L - 2)]) for gc in g]
    hsc = [sum(1 for p in range(n - 2) if p % 3 != 0 and s[p:p + 3] in STOPS) for s in seqs]
    gc3 = [sum(1 for p in range(2, n, 3) if s[p] in "GC") for s in seqs]
    hd = {(i, j): sum(a != b for a, b in zip(seqs[i], seqs[j])) for i in range(m) for j in range(i + 1, m)}
    return seqs, cai, cpb, hsc, gc3, hd


def test_eval_matches_closed_forms(cdp, orc):
    T = cdp.tables
    rng = np.random.default_rng(5)
    for trial in range(12):
        tables = T.synth_tables(trial) if trial % 2 else T.reference_tables()
        L = int(rng.integers(3, 60))
        m = int(rng.integers(2, 6))
        aa = T.synth_protein(L - 1, trial)
        pop = T.random_population(aa, m, 6, trial)
        obj, idx, pql = orc.evaluate(aa, tables, pop)
        for d in range(len(pop)):
            seqs, cai, cpb, hsc, gc3, hd = brute_objectives(cdp, aa, tables, pop[d])
            n = 3 * L
            assert abs(-obj[d, 0] - min(cai)) <= 1e-6 * min(cai)
            assert abs(-obj[d, 1] - min(cpb)) <= 1e-6 * max(1e-3, abs(min(cpb)))
            assert obj[d, 2] == -np.float32(min(hsc)) / np.float32(L - 1)
            assert idx[d, 4] == max(c for c in range(m) if hsc[c] == min(hsc))  # ties -> last CDS
            lo = min(hd.values())
            assert obj[d, 3] == -np.float32(lo) / np.float32(n)
            assert (idx[d, 6], idx[d, 7]) == max(k for k, v in hd.items() if v == lo)
            dev = [abs(np.float32(c) / np.float32(L) - tables.gc3_ref) for c in gc3]
            assert obj[d, 4] == max(dev)
            wc = max(c for c in range(m) if dev[c] == max(dev))
            assert idx[d, 8] == wc
            assert idx[d, 9] == (1 if np.float32(gc3[wc]) / np.float32(L) - tables.gc3_ref < 0 else -1)
            sl = [orc.sl(s, 2) for s in seqs]
            assert pql[d, 2] == max(sl)
            assert idx[d, 10] == max(c for c in range(m) if sl[c] == max(sl))
            assert obj[d, 5] == np.float32(max(sl)) / np.float32((n - 4) // 2)


def test_identical_copies_tie_to_last(cdp, orc, ref_tables):
    T = cdp.tables
    aa = T.synth_protein(40, 3)
    one = T.random_population(aa, 1, 1, 9)[0, 0]
    design = np.stack([one] * 4)[None]
    obj, idx, pql = orc.evaluate(aa, ref_tables, design)
    assert idx[0, 0] == idx[0, 2] == idx[0, 4] == idx[0, 8] == idx[0, 10] == 3
    assert (idx[0, 6], idx[0, 7]) == (2, 3)
    assert obj[0, 3] == 0.0


def test_reference_fp32_within_tolerance_of_fp64(cdp, orc, ref_tables):
    T = cdp.tables
    aa = T.synth_protein(500, 2)
    pop = T.random_population(aa, 4, 8, 1)
    o32, o64 = orc.cai_cpb_reference(aa, ref_tables, pop, threads=256)
    obj, _, _ = orc.evaluate(aa, ref_tables, pop)
    assert np.all(np.abs(o32[..., 0] / o64[..., 0] - 1) < 1e-4)  # the reference's own rounding level
    # canonical (engine) values: CAI within 1e-12, CPB within 1e-9 of FP64
    assert np.allclose(-obj[:, 0], o64[..., 0].min(axis=1), rtol=1e-7, atol=0)
    assert np.allclose(-obj[:, 1], o64[..., 1].min(axis=1), rtol=1e-6, atol=1e-9)


def hairpin(stem, loop, left="", right=""):
    return left + stem + loop + "".join(COMP[c] for c in reversed(stem)) + right


@pytest.mark.parametrize("k", [1, 3, 6, 12])
def test_hand_built_hairpins(orc, k):
    # a perfect k-pair stem closing a 4-nt loop has its innermost pair 5 apart -> l = k
    stem = ("GAU" * 8)[:k]
    assert orc.sl(hairpin(stem, "AAAA", "AAAA", "AAAA"), 2) == k
    assert orc.sl(hairpin(stem, "AAAA", "AAAA", "AAAA"), 0)[0] == k
    # a 3-nt loop: the innermost pair is only 4 apart and is excluded -> l = k-1
    assert orc.sl(hairpin(stem, "AAA", "AAAA", "AAAA"), 0)[0] == k - 1
    # the stem's position is reported as (p, q): first base of the left arm, innermost base of the right arm
    l, p, q = orc.sl(hairpin(stem, "AAAA", "AAAA", "AAAA"), 0)
    if k > 1:
        assert (l, p, q) == (k, 4, 4 + k + 4)


def test_no_complement_means_zero(orc):
    assert orc.sl("AAAAGGGGAAGAGAGAAA", 0) == (0, 0, 0)


def test_stem_closed_form_equals_walk_and_brute(orc):
    rng = np.random.default_rng(11)
    for _ in range(800):
        n = int(rng.integers(5, 70))
        probs = rng.dirichlet(np.ones(4))
        s = rng.choice(4, size=n, p=probs).astype(np.uint8)
        l0 = orc.sl(s, 0)[0]
        assert l0 == orc.sl(s, 1) == orc.sl(s, 2)


def test_b4twr7_denominators(cdp):
    """The reference's own output quantises onto integer counts over L-1,
    3L, (3L-4)/2 and L with L = 505 residues + appended stop (main.cu:518)."""
    aa = cdp.tables.parse_fasta(os.path.join(GOLD, "fasta", "B4TWR7.fasta.txt"))
    L = len(aa)
    assert L == 506
    rows = np.loadtxt(os.path.join(GOLD, "Normalized_value_quality_computation.txt"))
    assert rows.shape == (200, 6)
    # normalisation (main.cu:23-24, 820-835): v = (x - ideal) / (nadir - ideal)
    hsc = 0.4 * (1 - rows[:, 2]) * (L - 1)
    hd = 0.5 * (1 - rows[:, 3]) * (3 * L)
    sl = rows[:, 5] * ((3 * L - 4) // 2)
    for v in (hsc, hd, sl):
        assert np.max(np.abs(v - np.round(v))) < 2e-3
    gdev = 0.6 * rows[:, 4]
    cnt = np.minimum(np.abs(L * (0.381 + gdev) - np.round(L * (0.381 + gdev))),
                     np.abs(L * (0.381 - gdev) - np.round(L * (0.381 - gdev))))
    assert np.max(cnt) < 2e-3
    # a wrong L (no stop) breaks the quantisation
    bad = 0.4 * (1 - rows[:, 2]) * (L - 2)
    assert np.max(np.abs(bad - np.round(bad))) > 0.05
