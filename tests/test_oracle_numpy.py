"""Pins the numpy oracle (oracle/ref_numpy.py) by checks the mathematics fixes.

Each check is chosen so that a plausible mistake (dropped term, wrong sign,
wrong index or transposed operand) fails it:
  * grid: closed form of the tanh map, ghost-centre rule, dx = dxmin
  * kxx/kzz: eigenvalue identity of the periodic 3-point Laplacian
  * Dyy: W*Dyy symmetric (W = diag of cell heights), zero row sums
    (Neumann), and on a uniform grid the cosine-basis eigenvalues
    -(2-2cos(pi k/n))/h^2 (the closed form of the reference's
    tests/vof/spike_proj_lib.f90:26-32)
  * zgtsv: dense solves, including systems that force row interchanges
  * FFT stages: naive O(n^2) DFT
  * whole solve: dense solve of the assembled pinned 7-point operator
  * projection: manufactured u* = curl_h(A) + grad_h(phi_true) returns
    phi_true and curl_h(A); idempotence; divergence-free result
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import ref_numpy as R


def test_grid_closed_form_and_ghosts():
    g = R.channel_grid(12, 10, 8, 2.0)
    eta = np.arange(g.ny) / (g.ny - 1)
    y = g.Ly_i * 0.5 * (1 + np.tanh(2.0 * (eta - 0.5)) / math.tanh(1.0))
    assert np.allclose(g.y, y, rtol=0, atol=1e-15)
    assert g.y[0] == 0.0 and abs(g.y[-1] - 2.0) < 1e-15
    # centres and one mirrored ghost per side (initialization.f90:249-272)
    assert np.allclose(g.yg[1:-1], 0.5 * (g.y[:-1] + g.y[1:]), rtol=0, atol=0)
    assert g.yg[0] == pytest.approx(2 * g.y[0] - g.yg[1], abs=1e-15)
    assert g.yg[-1] == pytest.approx(2 * g.y[-1] - g.yg[-2], abs=1e-15)
    # periodic period Lx_i - dx = 2 pi, dx = dxmin
    assert g.dx == pytest.approx(2 * math.pi / 12, rel=1e-14)
    assert g.dz == pytest.approx(math.pi / 8, rel=1e-14)
    assert g.x[-1] - g.x[0] - g.dx == pytest.approx(2 * math.pi, rel=1e-14)


@pytest.mark.parametrize("n,d", [(8, 0.3), (12, 0.1), (18, 1.7), (1, 0.5)])
def test_wavenumbers_eigen_identity(n, d):
    k = R.wavenumbers(n, d)
    i = np.arange(n)
    for m in range(n):
        e = np.exp(2j * np.pi * m * i / n)
        lap = (np.roll(e, -1) - 2 * e + np.roll(e, 1)) / d ** 2
        assert np.allclose(lap, k[m] * e, atol=1e-12 * max(1.0, abs(k[m])))
    assert k[0] == 0.0
    # symmetric in +-m, bitwise
    assert all(k[m] == k[n - m] for m in range(1, n))


def test_dyy_symmetry_rowsums_and_uniform_limit():
    g = R.channel_grid(8, 13, 8, 2.5)
    D = R.dyy_matrix(g.y, g.yg)
    w = np.diff(g.y)  # cell heights
    WD = w[:, None] * D
    assert np.allclose(WD, WD.T, atol=1e-12 * np.abs(WD).max())
    assert np.allclose(D.sum(axis=1), 0.0, atol=1e-10 * np.abs(D).max())
    # tridiagonal
    assert np.count_nonzero(np.triu(D, 2)) == 0 and np.count_nonzero(np.tril(D, -2)) == 0
    # uniform grid: Neumann cosine basis eigenvalues
    gu = R.ChannelGrid(8, 16, 8, 1.0, 2.0, 1.0, 0.0, grid_type=1)
    Du = R.dyy_matrix(gu.y, gu.yg)
    h = 2.0 / 16
    lam = np.sort(np.linalg.eigvals(Du).real)
    ref = np.sort(-(2 - 2 * np.cos(np.pi * np.arange(16) / 16)) / h ** 2)
    assert np.allclose(lam, ref, atol=1e-9)


def test_zgtsv_against_dense_and_pivoting():
    rng = np.random.default_rng(1)
    M, n = 40, 9
    # general systems: row interchanges happen
    dl = rng.standard_normal((M, n - 1)); du = rng.standard_normal((M, n - 1)); d = rng.standard_normal((M, n))
    b = rng.standard_normal((M, n)) + 1j * rng.standard_normal((M, n))
    x, info = R.zgtsv(dl, d, du, b)
    assert np.all(info == 0)
    npiv = 0
    for m in range(M):
        T = np.diag(d[m]) + np.diag(dl[m], -1) + np.diag(du[m], 1)
        xr = np.linalg.solve(T, b[m])
        assert np.allclose(x[m], xr, rtol=1e-10, atol=1e-10 * np.abs(xr).max())
        npiv += int(np.any(np.abs(d[m][:-1]) < np.abs(dl[m])))
    assert npiv > 0
    # a zero pivot is reported, not divided by
    x, info = R.zgtsv(np.zeros((1, 2)), np.array([[0.0, 1.0, 1.0]]), np.zeros((1, 2)), np.ones((1, 3)))
    assert info[0] == 1


def test_fft_stage_matches_naive_dft():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((6, 12)) + 1j * rng.standard_normal((6, 12))
    assert np.allclose(np.fft.fft(a, axis=1), R.naive_dft(a, 1, -1), atol=1e-12)
    assert np.allclose(np.fft.ifft(a, axis=1) * 12, R.naive_dft(a, 1, +1), atol=1e-12)


def test_dyy_of_one_cell_between_the_walls():
    # initialization.f90:523-535 assigns Dyy(2,2) twice when nyg-1 = 2: b+c,
    # then a+b, which stays.  a+b = -c = -1/((y(2)-y(1)) (yg(2)-yg(1))).
    g = R.channel_grid(8, 1, 4, 1.5)
    D = R.dyy_matrix(g.y, g.yg)
    assert D.shape == (1, 1)
    c = 1.0 / ((g.y[1] - g.y[0]) * (g.yg[1] - g.yg[0]))
    assert D[0, 0] == pytest.approx(-c, rel=1e-14)


@pytest.mark.parametrize("dims", [(8, 6, 4), (6, 5, 6), (4, 3, 9), (8, 1, 6)])
def test_solve_matches_dense_pinned_operator(dims):
    nxp, nym, nzp = dims
    g = R.channel_grid(nxp, nym, nzp, 1.7)
    pr = R.Projection(g)
    rng = np.random.default_rng(3)
    _, _, _, _, rhs = R.new_fields(g)
    r = rng.standard_normal((nxp, nym, nzp))
    rhs[:nxp, :nym, :nzp] = r
    pr.solve_poisson_equation(rhs)
    A = R.dense_pinned_operator(g)
    x = np.linalg.solve(A, r.reshape(-1, order="F")).reshape((nxp, nym, nzp), order="F")
    assert np.max(np.abs(rhs[:nxp, :nym, :nzp] - x)) <= 1e-12 * np.max(np.abs(x))
    # periodic duplicates of projection.f90:455, :481
    assert np.array_equal(rhs[g.nxg - 3, :, :], rhs[0, :, :])
    assert np.array_equal(rhs[: g.nxg - 2, :, g.nzp], rhs[: g.nxg - 2, :, 0])


def _fill_periodic(g, U, V, W):
    U[0] = U[g.nx - 2]; U[g.nx - 1] = U[1]
    for F in (V, W):
        F[0] = F[g.nxg - 3]; F[g.nxg - 2] = F[1]; F[g.nxg - 1] = F[2]
    for F in (U, V):
        F[:, :, 0] = F[:, :, g.nzg - 3]; F[:, :, g.nzg - 2] = F[:, :, 1]; F[:, :, g.nzg - 1] = F[:, :, 2]
    W[:, :, 0] = W[:, :, g.nz - 2]; W[:, :, g.nz - 1] = W[:, :, 1]


def manufactured(g, seed):
    """u* = curl_h(A) + grad_h(phi_true) on the MAC grid, with the discrete
    operators of the reference's divergence and correction."""
    rng = np.random.default_rng(seed)
    nxp, nym, nzp = g.nxp, g.nym, g.nzp
    y, yg, z, zg = g.y, g.yg, g.z, g.zg
    dx = g.dx
    # vector potential: Az at (x-face i, y-face jf, cell k), Ax at (cell i, y-face jf, z-face kf), zero on the walls
    Az = rng.standard_normal((nxp, nym + 1, nzp)); Az[:, 0] = 0; Az[:, nym] = 0
    Ax = rng.standard_normal((nxp, nym + 1, nzp)); Ax[:, 0] = 0; Ax[:, nym] = 0
    dyc = (y[1:] - y[:-1])[None, :, None]          # cell heights y(j)-y(j-1)
    dzc = (z[1:nzp + 1] - z[0:nzp])[None, None, :]       # z(k)-z(k-1) of cell k
    u = (Az[:, 1:, :] - Az[:, :-1, :]) / dyc
    v = -(Az - np.roll(Az, 1, axis=0)) / dx + (Ax - np.roll(Ax, 1, axis=2)) / dzc
    w = -(Ax[:, 1:, :] - Ax[:, :-1, :]) / dyc
    phi = rng.standard_normal((nxp, nym, nzp))
    phi -= phi[:, 0, :].mean()                       # the pin's gauge: zero x-z mean of the first row
    gu = (np.roll(phi, -1, axis=0) - phi) / dx
    dyg = (yg[2:nym + 1] - yg[1:nym])[None, :, None]  # yg(j+1)-yg(j), interior y-faces
    gv = np.zeros_like(v); gv[:, 1:nym, :] = (phi[:, 1:, :] - phi[:, :-1, :]) / dyg
    dzg = (zg[2:nzp + 2] - zg[1:nzp + 1])[None, None, :]
    gw = (np.roll(phi, -1, axis=2) - phi) / dzg
    U, V, W, P, rhs = R.new_fields(g)
    # 0-based cell i is the reference's cell i+2; U(i+2) (array index i+1) is its
    # right face, V(jf+1) (index jf) the y face jf, W(k+2) (index k+1) its top z face
    U[1:nxp + 1, 1:nym + 1, 1:nzp + 1] = u + gu
    V[1:nxp + 1, 0:nym + 1, 1:nzp + 1] = v + gv
    W[1:nxp + 1, 1:nym + 1, 1:nzp + 1] = w + gw
    _fill_periodic(g, U, V, W)
    return U, V, W, rhs, phi, (u, v, w)


@pytest.mark.parametrize("dims,alpha", [((8, 6, 8), 1.5), ((12, 9, 6), 2.2)])
def test_manufactured_solution_and_idempotence(dims, alpha):
    g = R.channel_grid(*dims, alpha)
    U, V, W, rhs, phi, (u, v, w) = manufactured(g, 7)
    pr = R.Projection(g)
    pr.compute_projection_step(U, V, W, rhs)
    nxp, nym, nzp = dims
    err_phi = np.max(np.abs(rhs[:nxp, :nym, :nzp] - phi)) / np.max(np.abs(phi))
    assert err_phi < 1e-11, err_phi
    ucorr = U[1:nxp + 1, 1:nym + 1, 1:nzp + 1]
    assert np.max(np.abs(ucorr - u)) < 1e-10 * np.max(np.abs(u))
    assert np.max(np.abs(V[1:nxp + 1, 0:nym + 1, 1:nzp + 1] - v)) < 1e-10 * np.max(np.abs(v))
    assert np.max(np.abs(W[1:nxp + 1, 1:nym + 1, 1:nzp + 1] - w)) < 1e-10 * np.max(np.abs(w))
    # the corrected field is divergence free (ghosts refreshed as after apply_boundary_conditions)
    _fill_periodic(g, U, V, W)
    assert pr.check_divergence(U, V, W) < 1e-11 * max(np.abs(U).max(), np.abs(V).max()) / g.dx
    # projecting again changes nothing
    U2, V2, W2 = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    pr.compute_projection_step(U2, V2, W2, rhs)
    assert np.max(np.abs(rhs[:nxp, :nym, :nzp])) < 1e-11 * np.max(np.abs(phi))
    assert np.max(np.abs(U2 - U)) < 1e-13 * np.abs(U).max() * 100


def test_pressure_ghosts_follow_periodic_and_wall_rules():
    g = R.channel_grid(6, 5, 4, 1.5)
    pr = R.Projection(g)
    rng = np.random.default_rng(4)
    _, _, _, P, rhs = R.new_fields(g)
    rhs[: g.nxp, : g.nym, : g.nzp] = rng.standard_normal((g.nxp, g.nym, g.nzp))
    pr.solve_poisson_equation(rhs, P, dt=1e-3, rk_coef33=0.75)
    phi = rhs[: g.nxp, : g.nym, : g.nzp]
    for i in range(1, g.nxg + 1):
        for j in range(1, g.nyg + 1):
            for k in range(1, g.nzg + 1):
                ii = (i - 2) % g.nxp
                jj = min(max(j - 2, 0), g.nym - 1)
                kk = (k - 2) % g.nzp
                assert P[i - 1, j - 1, k - 1] == phi[ii, jj, kk] / (1e-3 * 0.75)
