"""Codon tables, protein input and synthetic workloads.

The genome encoding is the reference's (m CDS copies of one protein, each
codon a synonym of its amino acid) stored as int8 synonym indices instead of
ASCII nucleotides (include/utils.cuh:1072-1077, include/info.cuh:9-31).
"""
import json
import os
from dataclasses import dataclass

import numpy as np

# include/info.cuh:8-10
AMINO = "ACDEFGHIKLMNPQRSTVWYZ"
SYN_START = (0, 4, 6, 8, 10, 12, 16, 18, 21, 23, 29, 30, 32, 36, 38, 44, 50, 54, 58, 59, 61)
SYN_NUM = (4, 2, 2, 2, 2, 4, 2, 3, 2, 6, 1, 2, 4, 2, 6, 6, 4, 4, 1, 2, 3)
STOP = 20
NOBJ = 6

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "reference_tables.json")

# Amino-acid composition of UniProtKB/Swiss-Prot (percent), the "natural
# background frequencies" of the synthetic proteins.
BACKGROUND = {
    "A": 8.25, "R": 5.53, "N": 4.06, "D": 5.46, "C": 1.38, "Q": 3.93, "E": 6.72,
    "G": 7.07, "H": 2.27, "I": 5.91, "L": 9.65, "K": 5.80, "M": 2.41, "F": 3.86,
    "P": 4.74, "S": 6.63, "T": 5.35, "W": 1.10, "Y": 2.92, "V": 6.86,
}


@dataclass
class Tables:
    """Per-job codon table.

    codons  -- 192 ASCII letters (A,C,G,U): codon g is codons[3g:3g+3]; amino-acid
               blocks follow SYN_START/SYN_NUM, synonyms by ascending weight,
               the three stop codons last.
    weights -- float32[64] CAI weights in (0, 1].
    cps     -- float32[61*61] codon-pair scores, cps[a*61+b].
    gc3_ref -- reference GC3 fraction.
    """
    codons: str
    weights: np.ndarray
    cps: np.ndarray
    gc3_ref: float = 0.381

    def codon_bytes(self) -> bytes:
        return self.codons.encode("ascii")


def reference_tables() -> Tables:
    """The reference's own table (include/info.cuh:7-54)."""
    d = json.load(open(_DATA))
    return Tables(d["codons"], np.asarray(d["weights"], np.float32), np.asarray(d["cps"], np.float32),
                  float(np.float32(d["gc3_ref"])))


def synonym_sets():
    """The codon set of every amino acid (the standard genetic code)."""
    t = reference_tables()
    return [sorted(t.codons[3 * g:3 * g + 3] for g in range(SYN_START[a], SYN_START[a] + SYN_NUM[a]))
            for a in range(21)]


def amino_index(letter: str) -> int:
    """findAminoIndex (include/utils.cuh:60-86); unknown letters raise."""
    i = AMINO.find(letter)
    if i < 0 or len(letter) != 1:
        raise ValueError(f"unknown amino-acid letter {letter!r}")
    return i


def parse_fasta(path: str) -> np.ndarray:
    """Amino-acid indices of a FASTA file with the reference's semantics.

    src/main.cu:497-523 skips the first line, keeps every other byte except
    newlines and overwrites the final (EOF) byte with 'Z', so L = residues + 1
    and the last codon is a stop.  Letters outside the alphabet are rejected
    (the reference exits, utils.cuh:84-85); 'Z' may only be that final stop.
    """
    with open(path, "rb") as f:
        data = f.read()
    nl = data.find(b"\n")
    body = data[nl + 1:] if nl >= 0 else b""
    residues = [chr(c) for c in body if c != 0x0A]
    idx = [amino_index(c) for c in residues]
    if STOP in idx:
        raise ValueError("stop symbol 'Z' inside the protein")
    return np.asarray(idx + [STOP], np.int8)


def synth_protein(n_residues: int, seed: int) -> np.ndarray:
    """Met followed by n_residues-1 residues drawn i.i.d. from BACKGROUND, plus the stop."""
    rng = np.random.default_rng([seed, 0])
    letters = AMINO[:20]
    p = np.array([BACKGROUND[c] for c in letters], np.float64)
    p /= p.sum()
    body = rng.choice(20, size=n_residues - 1, p=p)
    return np.concatenate([[AMINO.index("M")], body, [STOP]]).astype(np.int8)


def synth_tables(seed: int) -> Tables:
    """Synthetic CAI and pair tables.

    Each sense codon gets a weight drawn U(0.05, 1], renormalised so the best
    synonym of every amino acid has weight 1, and synonyms are ordered by
    ascending weight (the ordering the CAI operator relies on).  The stop block
    keeps the reference's order and weights (198/549, 223/549, 1).  The pair
    table is 61x61 N(0, 0.3) in float32.
    """
    rng = np.random.default_rng([seed, 1])
    ref = reference_tables()
    codons = []
    weights = np.zeros(64, np.float32)
    for a in range(20):
        s0, ns = SYN_START[a], SYN_NUM[a]
        block = [ref.codons[3 * g:3 * g + 3] for g in range(s0, s0 + ns)]
        u = (1.0 - rng.uniform(0.0, 0.95, size=ns)).astype(np.float32)
        w = (u / u.max()).astype(np.float32)
        order = np.argsort(w, kind="stable")
        for j, o in enumerate(order):
            codons.append(block[o])
            weights[s0 + j] = w[o]
    for g in range(61, 64):
        codons.append(ref.codons[3 * g:3 * g + 3])
        weights[g] = ref.weights[g]
    cps = rng.normal(0.0, 0.3, size=61 * 61).astype(np.float32)
    return Tables("".join(codons), weights, cps, ref.gc3_ref)


def random_population(aa: np.ndarray, m: int, count: int, seed: int) -> np.ndarray:
    """Uniform random synonym indices [count][m][L] (int8) for tests and benches."""
    rng = np.random.default_rng([seed, 2])
    L = len(aa)
    ns = np.array([SYN_NUM[a] for a in aa], np.int64)
    u = rng.random((count, m, L))
    return np.floor(u * ns[None, None, :]).astype(np.int8)


def decode(aa: np.ndarray, tables: Tables, syn: np.ndarray) -> str:
    """Nucleotide string of one CDS from synonym indices."""
    return "".join(tables.codons[3 * (SYN_START[a] + int(s)):3 * (SYN_START[a] + int(s)) + 3]
                   for a, s in zip(aa, syn))


def encode(aa: np.ndarray, tables: Tables, cds: str) -> np.ndarray:
    """Synonym indices of one CDS given as nucleotides (inverse of decode)."""
    out = []
    for k, a in enumerate(aa):
        cod = cds[3 * k:3 * k + 3]
        block = [tables.codons[3 * g:3 * g + 3] for g in range(SYN_START[a], SYN_START[a] + SYN_NUM[a])]
        out.append(block.index(cod))
    return np.asarray(out, np.int8)
