"""Pins of the oracle's primitives: Philox4x32-10, exp, Das-Dennis, FASTA input."""
import itertools
import math
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_philox_known_answers(orc):
    # Known-answer vectors published with Random123 (Salmon et al., SC'11)
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, out in kat:
        assert list(orc.philox(ctr, key)) == out


def test_det_exp_accuracy(orc):
    for x in np.concatenate([np.linspace(-700, 0, 2001), -np.logspace(-12, 1, 300), [0.0, -0.5, -1e-300]]):
        got, want = orc.det_exp(x), math.exp(x)
        assert abs(got - want) <= 4e-16 * want + 1e-300, x


@pytest.mark.parametrize("H", [1, 2, 4, 5, 7])
def test_dasdennis_lattice_and_order(orc, H):
    pts = orc.dasdennis(6, H)
    assert pts.shape == (math.comb(6 + H - 1, H), 6)
    want = []
    for comb in itertools.combinations_with_replacement(range(6), H):
        v = [0] * 6
        for c in comb:
            v[c] += 1
        want.append(v)
    # coordinates are k repeated float32 additions of 1/H (NUnrank, sorting.cuh:1332)
    step = np.float32(1.0) / np.float32(H)
    acc = [np.float32(0)]
    for _ in range(H):
        acc.append(np.float32(acc[-1] + step))
    want_f = np.array([[acc[k] for k in v] for v in want], np.float32)
    assert np.array_equal(pts, want_f)


RESIDUES = {"A4Y1B6": 716, "A6L9J9": 221, "B3LS90": 679, "B4TWR7": 505, "B7KHU9": 38, "Q5VZP5": 1158,
            "Q88X33": 114, "Q89BP2": 388, "Q91X51": 446}


@pytest.mark.parametrize("name", sorted(RESIDUES))
def test_fasta_appends_stop(cdp, name):
    aa = cdp.tables.parse_fasta(os.path.join(GOLD, "fasta", f"{name}.fasta.txt"))
    assert len(aa) == RESIDUES[name] + 1  # main.cu:510-520: L = residues + 1
    assert aa[-1] == 20 and (aa[:-1] != 20).all()
    assert aa[0] == cdp.tables.AMINO.index("M")


def test_fasta_rejects_unknown(cdp, tmp_path):
    p = tmp_path / "bad.fasta"
    p.write_bytes(b">x\nMKXL\n")
    with pytest.raises(ValueError):
        cdp.tables.parse_fasta(str(p))


def test_reference_tables_are_the_standard_code(cdp, ref_tables):
    T = cdp.tables
    code = {}
    bases = "UCAG"
    aas = "FFLLSSSSYYZZCCZWLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG"
    for i, c in enumerate(itertools.product(bases, repeat=3)):
        code["".join(c)] = aas[i]
    for a in range(21):
        for g in range(T.SYN_START[a], T.SYN_START[a] + T.SYN_NUM[a]):
            assert code[ref_tables.codons[3 * g:3 * g + 3]] == T.AMINO[a]
    assert sorted(ref_tables.codons[183:192][i:i + 3] for i in (0, 3, 6)) == ["UAA", "UAG", "UGA"]


def test_synth_tables_keep_the_code(cdp):
    T = cdp.tables
    sets = T.synonym_sets()
    for seed in (1, 2, 3):
        t = T.synth_tables(seed)
        for a in range(21):
            blk = [t.codons[3 * g:3 * g + 3] for g in range(T.SYN_START[a], T.SYN_START[a] + T.SYN_NUM[a])]
            assert sorted(blk) == sets[a]
            w = t.weights[T.SYN_START[a]:T.SYN_START[a] + T.SYN_NUM[a]]
            assert (np.diff(w) >= 0).all() and w[-1] == 1.0 and (w > 0.05).all()
        assert t.cps.dtype == np.float32 and abs(float(t.cps.std()) - 0.3) < 0.02
