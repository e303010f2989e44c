"""Pins of the oracle's generation loop (src/main.cu:121-240, 700-732).

Each child's mutation type is recomputed from its Philox stream and the child
is checked against what that operator (include/mutation.cuh) may change.
"""
import numpy as np
import pytest

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
