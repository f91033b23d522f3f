"""Shared cases and helpers of the parity tests."""
from __future__ import annotations

import math

import numpy as np

import dfp
from oracle import ref_c

# (name, grid builder args, sigma): sizes are Fourier/cell counts nxp x nym x nzp
SMALL_CASES = [
    ("tiny_8x6x4", dict(nxp=8, nym=6, nzp=4, alpha=1.5), 0.1),
    ("asym_12x10x8", dict(nxp=12, nym=10, nzp=8, alpha=2.0), 0.1),
    ("radix3_24x9x18", dict(nxp=24, nym=9, nzp=18, alpha=1.2), 0.1),
    ("odd_nym_32x33x48", dict(nxp=32, nym=33, nzp=48, alpha=2.5), 0.15),
    ("flat_uniform_16x2x16", dict(nxp=16, nym=2, nzp=16, alpha=0.0, grid_type=1), 0.1),
    ("nzp1_6x7x1", dict(nxp=6, nym=7, nzp=1, alpha=1.0), 0.1),
    ("one_cell_8x1x12", dict(nxp=8, nym=1, nzp=12, alpha=1.5), 0.1),
]


def make_grid(nxp, nym, nzp, alpha, grid_type=2, **kw):
    return dfp.channel_grid(nxp, nym, nzp, alpha, grid_type=grid_type, **kw)


def chan_les_small_grid():
    # tests/regression/chan_les_small/input_parameters:6-14 of the reference
    return dfp.namelist_grid(18, 17, 18, 3.0, 2.0, 1.5, 2.0)


def fields(g, seed, sigma):
    U, V, W = dfp.synth_fields(g, seed, sigma)
    P = np.zeros(g.shape_p, order="F")
    rhs = np.zeros(g.shape_rhs, order="F")
    return U, V, W, P, rhs


def rel_linf(a, b):
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


def oracle_project(g, U, V, W, P, rhs, dt, c, save_p=True, nthreads=0):
    op = ref_c.CProjection(g, nthreads)
    try:
        op.compute_projection_step(U, V, W, rhs, P if save_p else None, dt, c)
    finally:
        op.close()


def div_bound(g, U, V, W):
    """1e-12 * max|u| / min spacing"""
    umax = max(np.max(np.abs(U)), np.max(np.abs(V)), np.max(np.abs(W)))
    return 1e-12 * umax / g.min_spacing
