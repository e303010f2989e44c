"""Synthetic job mixes for the batched and served forms.

Every job is a synthetic protein (Met + i.i.d. residues with background
frequencies, plus the stop) with its own synthetic CAI and pair tables, six
objectives and the Das-Dennis H given, as the single-job workloads are built.
Seeds are fixed so a mix is reproducible: job i of a mix with master seed s
uses seed 1000 * s + i for its protein, tables and Philox key.
"""
from dataclasses import dataclass

import numpy as np

from . import tables as T


@dataclass
class JobSpec:
    residues: int     # protein length without the stop (L = residues + 1)
    m: int            # CDS copies
    N: int            # population
    H: int            # Das-Dennis divisions
    generations: int
    seed: int
    Pm: float = 0.1

    def build(self):
        """(aa, tables, m, N, H, Pm, seed): the arguments of Job()."""
        return (T.synth_protein(self.residues, self.seed), T.synth_tables(self.seed), self.m, self.N, self.H,
                self.Pm, self.seed)


def batched_mix(jobs=64, master_seed=4, generations=500):
    """64 independent jobs: lengths U{50..2000} aa, m in {2,4,8}, N in
    {256,512,1024}, H = 4, 500 generations each."""
    rng = np.random.default_rng([master_seed, 3])
    out = []
    for i in range(jobs):
        out.append(JobSpec(residues=int(rng.integers(50, 2001)), m=int(rng.choice([2, 4, 8])),
                           N=int(rng.choice([256, 512, 1024])), H=4, generations=generations,
                           seed=1000 * master_seed + i))
    return out


@dataclass
class Arrival:
    t: float          # seconds after the start of the run
    spec: JobSpec


def served_arrivals(rate=20.0, duration=600.0, master_seed=5):
    """Open-loop Poisson arrivals: protein length lognormal (median 350 aa,
    sigma 0.8) clipped to [30, 3000], m in {2,4,8}, N uniform over the
    integers {256..2048}, generation budget uniform over {300..1000}, H = 4."""
    rng = np.random.default_rng([master_seed, 4])
    out, t, i = [], 0.0, 0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            return out
        res = int(np.clip(np.round(350.0 * np.exp(0.8 * rng.standard_normal())), 30, 3000))
        out.append(Arrival(t, JobSpec(residues=res, m=int(rng.choice([2, 4, 8])), N=int(rng.integers(256, 2049)),
                                      H=4, generations=int(rng.integers(300, 1001)),
                                      seed=1000 * master_seed + i)))
        i += 1
