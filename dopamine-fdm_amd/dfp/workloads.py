"""Synthetic workloads of the batched and serving forms (seeded, reproducible).

The reference has no counterpart: it projects one grid per run.  These
generators fix the case mixes the batched executor and the serving loop are
measured on: stretched channels (tanh stretching, Lx = 2 pi, Lz = pi) with
seeded Gaussian-perturbed Poiseuille fields.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

BATCH_XZ = (32, 48, 64, 96, 128)
BATCH_Y = (32, 64, 96)


@dataclass(frozen=True)
class CaseSpec:
    nxp: int
    nym: int
    nzp: int
    alpha: float
    seed: int
    sigma: float = 0.1

    @property
    def key(self):
        return (self.nxp, self.nym, self.nzp)


def batched_cases(n: int = 48, seed: int = 48) -> list[CaseSpec]:
    """n independent cases: nxp, nzp from {32,48,64,96,128}, nym from
    {32,64,96}, stretching uniform in [1.0, 2.5], per-case field seeds."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nxp = int(rng.choice(BATCH_XZ))
        nzp = int(rng.choice(BATCH_XZ))
        nym = int(rng.choice(BATCH_Y))
        alpha = float(rng.uniform(1.0, 2.5))
        out.append(CaseSpec(nxp, nym, nzp, alpha, 1000 + i))
    return out


def smooth_sizes(lo: int, hi: int) -> list[int]:
    """Numbers 2^a 3^b in [lo, hi]."""
    v = []
    a = 1
    while a <= hi:
        b = a
        while b <= hi:
            if b >= lo:
                v.append(b)
            b *= 3
        a *= 2
    return sorted(v)


def snap(d: float, choices: list[int]) -> int:
    """The 2^a 3^b value nearest d (in log distance)."""
    return min(choices, key=lambda c: (abs(math.log(c / d)), c))


def serving_size(s: float, dmin=(32, 32, 32), dmax=(256, 192, 256)) -> tuple[int, int, int]:
    """Size of a request with log-uniform parameter s in [0, 1]: each
    dimension d = dmin * (dmax/dmin)^s snapped to the nearest 2^a 3^b (even in x)."""
    dims = []
    for lo, hi in zip(dmin, dmax):
        d = lo * (hi / lo) ** s
        ch = smooth_sizes(lo, hi)
        dims.append(snap(d, ch))
    if dims[0] % 2:
        dims[0] = snap(dims[0], [c for c in smooth_sizes(dmin[0], dmax[0]) if c % 2 == 0])
    return tuple(dims)
