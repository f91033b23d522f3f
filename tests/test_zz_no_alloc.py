"""dfp_max_divergence_noalloc (the allocation-free variant of
dfp_max_divergence) through the C ABI: the same value, bit for bit, as the
plan-workspace entry point, on two calls in a row (the second on a smaller
grid, so the fixed workspace holds the first call's partials when it
starts), on a grid with more cells than 1024 blocks of 256 cover (each block
strides), and for the inflow/outflow x."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import dfp

pytestmark = pytest.mark.gpu


def _device_fields(g, seed, sigma):
    U, V, W = dfp.synth_fields(g, seed, sigma)
    # a few large values so the maximum sits in one cell, not in the noise
    rng = np.random.default_rng(seed)
    for F in (U, V, W):
        idx = tuple(int(rng.integers(1, s - 1)) for s in F.shape)
        F[idx] += 50.0 * (seed % 7 + 1)
    return [dfp.DeviceArray.from_host(a) for a in (U, V, W)]


def _both(plan, dev):
    L = dfp.lib()
    a, b = ctypes.c_double(), ctypes.c_double(-1.0)
    rc = L.dfp_max_divergence(plan.handle, dev[0].ptr, dev[1].ptr, dev[2].ptr, ctypes.byref(a),
                              dfp.DFP_FIELDS_ON_DEVICE)
    assert rc == 0
    rc = L.dfp_max_divergence_noalloc(plan.handle, dev[0].ptr, dev[1].ptr, dev[2].ptr, ctypes.byref(b),
                                      dfp.DFP_FIELDS_ON_DEVICE)
    assert rc == 0, dfp.lib().dfp_last_error(plan.handle)
    return a.value, b.value


@pytest.mark.parametrize("x_bc_type", [0, 1])
def test_noalloc_max_divergence_equals_plan_workspace(x_bc_type):
    # first a grid whose cells need more than 1024 blocks of 256 (blocks
    # stride), then a smaller one that reuses the workspace
    for dims, seed in (((128, 96, 64), 3), ((12, 10, 8), 4)):
        g = dfp.channel_grid(*dims, 1.6, x_bc_type=x_bc_type)
        print(f"noalloc {dims} x_bc_type {x_bc_type}", flush=True)
        plan = dfp.Plan(g)
        dev = _device_fields(g, seed, 0.2)
        a, b = _both(plan, dev)
        assert a > 0.0 and a == b, (dims, a, b)
        plan.close()


def test_noalloc_refuses_host_fields():
    g = dfp.channel_grid(8, 6, 4, 1.5)
    plan = dfp.Plan(g)
    U, V, W = dfp.synth_fields(g, 1, 0.1)
    out = ctypes.c_double()
    rc = dfp.lib().dfp_max_divergence_noalloc(plan.handle, U.ctypes.data, V.ctypes.data, W.ctypes.data,
                                              ctypes.byref(out), 0)
    assert rc == -1
    plan.close()
