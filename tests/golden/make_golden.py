#!/usr/bin/env python3
"""Writes the golden fixtures of the projection path with the numpy oracle.

Each fixture holds the inputs U, V, W (seeded synthetic channel fields) and
the outputs of one projection at RK stage 3 (P saved): corrected U, V, W,
rhs_p (= phi with its periodic duplicates) and P, plus the grid and the
parameters.  They pin the C oracle (which then pins the GPU path) and are
also compared with the GPU path directly.

    python tests/golden/make_golden.py      # rewrites tests/golden/*.npz
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dopamine-fdm_amd"))

from oracle import ref_numpy as R  # noqa: E402
from dfp.synth import synth_fields  # noqa: E402

DT = 1e-3
C3 = 3.0 / 4.0

# name: (namelist nx, ny, nz, Lx, Ly, Lz, alpha, grid_type, seed, sigma)
CASES = {
    # periodic lengths 2*pi and pi, i.e. namelist Lx = 2*pi*(nxp+1)/nxp
    "tiny_8x6x4": (10, 7, 6, 2 * np.pi * 9 / 8, 2.0, np.pi * 5 / 4, 1.5, 2, 11, 0.1),
    "asym_12x10x8": (14, 11, 10, 2 * np.pi * 13 / 12, 2.0, np.pi * 9 / 8, 2.0, 2, 12, 0.1),
    # the reference's own regression case tests/regression/chan_les_small/input_parameters:6-14
    "chan_les_small_16x16x16": (18, 17, 18, 3.0, 2.0, 1.5, 2.0, 2, 13, 0.1),
}


def make(name):
    nx, ny, nz, Lx, Ly, Lz, alpha, gt, seed, sigma = CASES[name]
    g = R.namelist_grid(nx, ny, nz, Lx, Ly, Lz, alpha, gt)
    U, V, W = synth_fields(g, seed, sigma)
    P = np.zeros(g.shape_p, order="F")
    rhs = np.zeros(g.shape_rhs, order="F")
    out = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    R.Projection(g).compute_projection_step(out[0], out[1], out[2], out[4], out[3], DT, C3)
    np.savez_compressed(
        os.path.join(HERE, f"{name}.npz"),
        namelist=np.array([nx, ny, nz], dtype=np.int64),
        lengths=np.array([Lx, Ly, Lz]), alpha=np.array(alpha), grid_type=np.array(gt),
        seed=np.array(seed), sigma=np.array(sigma), dt=np.array(DT), c=np.array(C3),
        U_in=U, V_in=V, W_in=W,
        U_out=out[0], V_out=out[1], W_out=out[2], P_out=out[3], rhs_out=out[4],
        x=g.x, y=g.y, z=g.z, yg=g.yg, zg=g.zg, dx=np.array(g.dx), dz=np.array(g.dz),
    )


if __name__ == "__main__":
    for n in CASES:
        make(n)
        print("wrote", n)
