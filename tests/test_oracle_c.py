"""Pins the C oracle (oracle/dfp_oracle.c, the checker of the GPU tests and the
host-CPU baseline) against the numpy oracle and the golden fixtures."""
from __future__ import annotations

import glob
import os

import numpy as np
import pytest

import dfp
from oracle import ref_c
from oracle import ref_numpy as R
from _cases import SMALL_CASES, fields, make_grid, rel_linf

# projection fixtures (the snapshot-reader fixture has its own test)
GOLDEN = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
                if not os.path.basename(p).startswith("snapshot_"))


def _np_grid(g):
    return R.ChannelGrid(g.nxp, g.nym, g.nzp, g.Lx_i, g.Ly_i, g.Lz_i, g.alpha, g.grid_type)


@pytest.mark.parametrize("name,kw,sigma", SMALL_CASES, ids=[c[0] for c in SMALL_CASES])
def test_c_oracle_matches_numpy_oracle(name, kw, sigma):
    g = make_grid(**kw)
    gn = _np_grid(g)
    U, V, W, P, rhs = fields(g, 77, sigma)
    a = [x.copy(order="F") for x in (U, V, W, P, rhs)]
    b = [x.copy(order="F") for x in (U, V, W, P, rhs)]
    R.Projection(gn).compute_projection_step(a[0], a[1], a[2], a[4], a[3], 1e-3, 0.75)
    op = ref_c.CProjection(g, 2)
    op.compute_projection_step(b[0], b[1], b[2], b[4], b[3], 1e-3, 0.75)
    for lab, x, y in zip("UVWP", b[:4], a[:4]):
        assert rel_linf(x, y) <= 1e-13, lab
    assert rel_linf(b[4], a[4]) <= 1e-13
    assert op.check_divergence(*b[:3]) == R.Projection(gn).check_divergence(*b[:3])


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_c_oracle_matches_golden(path):
    f = np.load(path)
    nx, ny, nz = (int(v) for v in f["namelist"])
    Lx, Ly, Lz = (float(v) for v in f["lengths"])
    g = dfp.namelist_grid(nx, ny, nz, Lx, Ly, Lz, float(f["alpha"]), int(f["grid_type"]))
    # the grid is rebuilt bit for bit
    for k in ("x", "y", "z", "yg", "zg"):
        assert np.array_equal(getattr(g, k), f[k]), k
    assert g.dx == float(f["dx"]) and g.dz == float(f["dz"])
    U, V, W = (np.asfortranarray(f[k]) for k in ("U_in", "V_in", "W_in"))
    # the synthetic generator reproduces the stored inputs exactly
    Us, Vs, Ws = dfp.synth_fields(g, int(f["seed"]), float(f["sigma"]))
    assert np.array_equal(Us, U) and np.array_equal(Vs, V) and np.array_equal(Ws, W)
    P = np.zeros(g.shape_p, order="F")
    rhs = np.zeros(g.shape_rhs, order="F")
    op = ref_c.CProjection(g, 1)
    op.compute_projection_step(U, V, W, rhs, P, float(f["dt"]), float(f["c"]))
    for lab, x, key in (("U", U, "U_out"), ("V", V, "V_out"), ("W", W, "W_out"), ("P", P, "P_out"), ("phi", rhs, "rhs_out")):
        assert rel_linf(x, f[key]) <= 1e-13, lab


def test_c_zgtsv_is_the_numpy_restatement():
    rng = np.random.default_rng(5)
    for _ in range(30):
        n = int(rng.integers(1, 12))
        dl = rng.standard_normal(max(n - 1, 0)); du = rng.standard_normal(max(n - 1, 0)); d = rng.standard_normal(n)
        b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        xc, ic = ref_c.zgtsv(dl, d, du, b)
        xn, inn = R.zgtsv(dl[None, :], d[None, :], du[None, :], b[None, :])
        assert ic == inn[0]
        # same algorithm and pivot decisions; numpy's complex kernels may round the
        # real-by-complex products differently in the last bit
        assert np.max(np.abs(xc - xn[0])) <= 4e-16 * np.max(np.abs(xn[0])) * n


@pytest.mark.parametrize("n", [1, 2, 3, 6, 8, 12, 16, 48, 64, 96, 128, 5, 7, 35])
def test_c_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    for stride, dist, howmany in ((1, n, 5), (20, 1, 20)):
        a = rng.standard_normal((howmany, n)) + 1j * rng.standard_normal((howmany, n))
        buf = np.zeros(max(stride * n, dist * howmany) + howmany * n, dtype=np.complex128)
        for b in range(howmany):
            buf[b * dist + np.arange(n) * stride] = a[b]
        for sign, ref in ((-1, np.fft.fft(a, axis=1)), (1, np.fft.ifft(a, axis=1) * n)):
            x = buf.copy()
            ref_c.lib().ora_fft_many(n, howmany, x.ctypes.data, stride, dist, sign)
            got = np.array([x[b * dist + np.arange(n) * stride] for b in range(howmany)])
            assert np.max(np.abs(got - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))


def test_c_operators_match_numpy():
    g = make_grid(24, 17, 18, 2.0)
    gn = _np_grid(g)
    op = ref_c.CProjection(g)
    kxx, kzz, dl, d, du = op.operators()
    assert np.array_equal(kxx, R.wavenumbers(g.nxp, g.dx))
    assert np.array_equal(kzz, R.wavenumbers(g.nzp, g.dz))
    D = R.dyy_matrix(gn.y, gn.yg)
    assert np.array_equal(d, np.diag(D)) and np.array_equal(dl, np.diag(D, -1)) and np.array_equal(du, np.diag(D, 1))
    arr = ref_c.channel_grid_arrays(g.nxp, g.nym, g.nzp, g.Lx_i, g.Ly_i, g.Lz_i, g.alpha, 2)
    for k in ("x", "y", "z", "yg", "zg"):
        assert np.array_equal(arr[k], getattr(g, k)), k
    assert arr["dx"] == g.dx and arr["dz"] == g.dz


def test_c_oracle_threads_do_not_change_results():
    g = make_grid(32, 20, 24, 1.8)
    U, V, W, P, rhs = fields(g, 3, 0.1)
    outs = []
    for nt in (1, 4):
        a = [x.copy(order="F") for x in (U, V, W, P, rhs)]
        ref_c.CProjection(g, nt).compute_projection_step(a[0], a[1], a[2], a[4], a[3], 1e-3, 0.75)
        outs.append(a)
    for x, y in zip(*outs):
        assert np.array_equal(x, y)
