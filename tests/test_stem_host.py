"""The engine's stem scans (csrc/cd_scan.h along anti-diagonals, csrc/cd_tile.h
across them, csrc/cd_qgram.h the seeded q-gram search), compiled for the host
with the address sanitizer and run lane by lane as the stem kernels run them,
against the oracle's closed form (with its tie rule) on random, repetitive,
skewed and planted-hairpin sequences, with and without a seed (a stem length
known to exist), with guesses of the parent's length that are right, too
small or too large, and with hit budgets that run out."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "protein_encoding_v2_nsga-iii_amd", "csrc")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMP = [3, 2, 1, 0]


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("stem") / "stem_host_check")
    subprocess.run([HIPCC, "-O1", "-std=c++17", "--offload-host-only", "-x", "hip", "-Xarch_host",
                    "-fsanitize=address", "-I", CSRC, os.path.join(HERE, "stem_host_check.cpp"), "-o", out],
                   check=True, capture_output=True)
    return out


def sequences(seed, count, nmax):
    rng = np.random.default_rng(seed)
    out = []
    for t in range(count):
        n = int(rng.integers(1, nmax))
        kind = t % 5
        if kind == 0:
            s = rng.integers(0, 4, n)
        elif kind == 1:  # short repeats with a few substitutions: long runs
            s = np.resize(rng.integers(0, 4, int(rng.integers(1, 7))), n)
            f = rng.random(n) < 0.03
            s[f] = rng.integers(0, 4, int(f.sum()))
        elif kind == 2:  # planted hairpin of 5..60 pairs
            s = rng.integers(0, 4, n)
            k = int(rng.integers(5, max(6, min(60, n // 3))))
            i = int(rng.integers(0, max(1, n - 2 * k - 8)))
            if i + 2 * k + 5 <= n:
                s[i + k + 5:i + 2 * k + 5] = [COMP[x] for x in s[i:i + k][::-1]]
        elif kind == 3:  # A/U rich: many equal-length stems, exercises the tie rule
            s = rng.choice(4, size=n, p=[0.45, 0.05, 0.05, 0.45])
        else:  # perfect palindrome: one stem spanning the sequence
            h = rng.integers(0, 4, max(1, n // 2))
            s = np.concatenate([h, rng.integers(0, 4, n % 2 + 6), [COMP[x] for x in h[::-1]]])
        out.append(np.asarray(s, np.uint8))
    return out


def run(harness, seqs, seeds, lanes, parts=1, mode="diag", parent=None, budget=None):
    parent = parent or [0] * len(seqs)
    inp = "".join("".join(map(str, s)) + f" {k} {g}\n" for s, k, g in zip(seqs, seeds, parent))
    args = [harness, str(lanes), str(parts), mode] + ([str(budget)] if budget is not None else [])
    r = subprocess.run(args, input=inp, capture_output=True, text=True, check=True)
    got = [tuple(map(int, o.split())) for o in r.stdout.strip().split("\n")]
    if mode == "qgram":  # parts that finished, missed a guessed K, ran out of budget
        f = r.stderr.split()
        return got, {f[k]: int(f[k + 1]) for k in range(0, 6, 2)}
    return got


@pytest.mark.parametrize("mode,lanes,parts", [("diag", 256, 4), ("diag", 256, 1), ("diag", 1, 1), ("diag", 7, 3),
                                              ("diag", 64, 2), ("diag", 1024, 1), ("tile", 256, 1), ("tile", 64, 1),
                                              ("tile", 256, 4), ("tile", 7, 1), ("tile", 1, 1),
                                              ("tile", 7, 3), ("tile", 64, 2)])
def test_stem_scan_matches_oracle(orc, harness, mode, lanes, parts):
    seqs = [s for s in sequences(lanes + parts, 200, 1600) if len(s) > 0]
    want = [orc.sl(s, 0) for s in seqs]
    rng = np.random.default_rng(lanes + 7)
    for seeds in ([0] * len(seqs),
                  [w[0] for w in want],                                     # the exact length
                  [int(rng.integers(0, w[0] + 1)) for w in want]):          # any lower bound
        got = run(harness, seqs, seeds, lanes, parts, mode)
        for s, g, w, k in zip(seqs, got, want, seeds):
            assert g == w, (len(s), k)


@pytest.mark.parametrize("mode", ["diag", "tile"])
def test_stem_scan_edge_lengths(orc, harness, mode):
    # every length around the 32-position chunk boundaries and the hairpin limit
    rng = np.random.default_rng(3)
    seqs = []
    for n in list(range(1, 80)) + [95, 96, 97, 127, 128, 129, 1503, 4500]:
        seqs.append(np.asarray(rng.integers(0, 4, n), np.uint8))
        seqs.append(np.full(n, 0, np.uint8) if n % 2 else np.asarray([0, 3] * (n // 2), np.uint8))
    want = [orc.sl(s, 0) for s in seqs]
    got = run(harness, seqs, [0] * len(seqs), 256, 4 if mode == "diag" else 2, mode)
    for s, g, w in zip(seqs, got, want):
        assert g == w, len(s)


@pytest.mark.parametrize("lanes", [256, 64, 7, 1])
def test_qgram_search_matches_oracle(orc, harness, lanes):
    # a changed CDS as the offspring stem kernel takes it: seeds that are lower
    # bounds (exact search when 7 or more) and parent lengths that are right,
    # too small or too large (guessed K, then 7: exact when found, tiles when
    # not)
    seqs = [s for s in sequences(lanes * 3 + 1, 160, 2400) if len(s) > 0]
    rng = np.random.default_rng(lanes + 11)
    seqs += [np.asarray(rng.integers(0, 4, n), np.uint8) for n in (1503, 3000, 4503)]
    want = [orc.sl(s, 0) for s in seqs]
    cases = (([w[0] for w in want], [w[0] for w in want]),
             ([int(rng.integers(0, w[0] + 1)) for w in want], [w[0] + int(rng.integers(-3, 6)) for w in want]),
             ([min(w[0], int(rng.integers(0, 7))) for w in want], [w[0] + 2 for w in want]),
             ([min(w[0], int(rng.integers(0, 7))) for w in want], [w[0] + 3 for w in want]),
             ([min(w[0], int(rng.integers(0, 7))) for w in want], [0] * len(want)))
    paths = []
    for seeds, parent in cases:
        got, path = run(harness, seqs, seeds, lanes, 1, "qgram", parent)
        for s, g, w, k, pl in zip(seqs, got, want, seeds, parent):
            assert g == w, (len(s), k, pl)
        paths.append(path)
    assert all(p["qgram"] > 0 for p in paths), paths  # every case found stems by q-grams
    assert paths[3]["miss"] > 0, paths  # a guess above the longest stem misses, then 7 is searched


@pytest.mark.parametrize("budget", [0, 1, 3])
def test_qgram_search_budget(orc, harness, budget):
    # searches that run out of hit budget leave the CDS to the tiles
    seqs = [s for s in sequences(budget + 40, 120, 1600) if len(s) > 0]
    want = [orc.sl(s, 0) for s in seqs]
    for seeds, parent in (([w[0] for w in want], [0] * len(want)), ([0] * len(want), [w[0] for w in want])):
        got, path = run(harness, seqs, seeds, 64, 1, "qgram", parent, budget)
        for s, g, w in zip(seqs, got, want):
            assert g == w, len(s)
        assert path["overflow"] > 0, path
