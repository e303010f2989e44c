"""The engine's bit-sliced stem scan (csrc/cd_stem.h), compiled for the host
and run lane by lane, against the oracle's closed form (with its tie rule) on
random, repetitive, skewed and planted-hairpin sequences.  Stems shorter
than the smallest window (4) are out of the bit-sliced pass's scope: the
kernel rescans those CDS exactly, which the GPU parity tests cover."""
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
    subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-host-only", "-x", "hip", "-I", CSRC,
                    os.path.join(HERE, "stem_host_check.cpp"), "-o", out], check=True, capture_output=True)
    return out


def sequences(seed, count, nmax):
    rng = np.random.default_rng(seed)
    out = []
    for t in range(count):
        n = int(rng.integers(6, nmax))
        kind = t % 4
        if kind == 0:
            s = rng.integers(0, 4, n)
        elif kind == 1:
            s = np.resize(rng.integers(0, 4, int(rng.integers(1, 7))), n)
            f = rng.random(n) < 0.03
            s[f] = rng.integers(0, 4, int(f.sum()))
        elif kind == 2:
            s = rng.integers(0, 4, n)
            k = int(rng.integers(5, max(6, min(60, n // 3))))
            i = int(rng.integers(0, max(1, n - 2 * k - 8)))
            if i + 2 * k + 5 <= n:
                s[i + k + 5:i + 2 * k + 5] = [COMP[x] for x in s[i:i + k][::-1]]
        else:
            s = rng.choice(4, size=n, p=[0.45, 0.05, 0.05, 0.45])
        out.append(np.asarray(s, np.uint8))
    return out


@pytest.mark.parametrize("lanes,seed_t", [(256, 0), (1, 0), (7, 6), (1024, 0)])
def test_bit_sliced_stem_matches_oracle(orc, harness, lanes, seed_t):
    seqs = sequences(lanes + seed_t, 160, 1300)
    inp = "\n".join("".join(map(str, s)) for s in seqs) + "\n"
    out = subprocess.run([harness, str(lanes), str(seed_t)], input=inp, capture_output=True, text=True,
                         check=True).stdout.split("\n")
    checked = 0
    for s, o in zip(seqs, out):
        want = orc.sl(s, 0)
        if want[0] >= max(4, seed_t):
            assert tuple(map(int, o.split())) == want, len(s)
            checked += 1
    assert checked > 100
