"""Pins the x inflow/outflow branch of the oracles (x_bc_type = 1: DCT-IV in
x, src/projection.f90:126-129, :333-339; backflow term :65-96; outflow face
:578-589) by checks the mathematics fixes:
  * the DCT-IV (FFTW REDFT11): a naive O(n^2) sum, for scipy's stand-in in the
    numpy oracle and for the C oracle's half-length-FFT algorithm (odd n too)
  * kxx (initialization.f90:443-445): eigenvalues of the 3-point Laplacian
    with a Neumann inflow ghost and a zero outflow face value, on the DCT-IV
    basis
  * the solve: a dense solve of the assembled 7-point operator (no pin)
  * the backflow face value: u* = 0 with a uniform face value c must give
    phi = c everywhere (the constant satisfies every discrete equation), which
    fixes the sign and the 2/dx^2 weight of the right-hand-side term
  * the projection: a manufactured u* = curl_h(A) + grad_h(phi_true) returns
    phi_true and curl_h(A); idempotence; a divergence-free result with the
    backflow term on
  * P ghosts: Neumann copies in x, the wall copies in y, periodic z
  * the C oracle equals the numpy oracle (1e-13) with and without backflow
"""
from __future__ import annotations

import math

import numpy as np
import pytest
import scipy.fft

from oracle import ref_c
from oracle import ref_numpy as R


def io_grid(nxp, nym, nzp, alpha=1.7):
    return R.channel_grid(nxp, nym, nzp, alpha, x_bc_type=1)


@pytest.mark.parametrize("n", [1, 2, 7, 8, 12, 18, 96])
def test_dct4_naive_scipy_and_c(n):
    a = np.random.default_rng(n).standard_normal((4, n))
    ref = R.naive_dct4(a, 1)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(scipy.fft.dct(a, type=4, axis=1) - ref)) <= 1e-14 * scale
    assert np.max(np.abs(ref_c.dct4(a, 1) - ref)) <= 1e-14 * scale
    # self-inverse up to 2n
    assert np.allclose(R.naive_dct4(ref, 1) / (2 * n), a, rtol=0, atol=1e-13)


@pytest.mark.parametrize("n,d", [(8, 0.3), (12, 0.1), (5, 1.7), (1, 0.5)])
def test_dct_wavenumbers_eigen_identity(n, d):
    k = R.wavenumbers_dct(n, d)
    # 3-point Laplacian: phi_{-1} = phi_0 (Neumann), phi_n = -phi_{n-1} (zero face value)
    L = np.zeros((n, n))
    for i in range(n):
        L[i, i] -= 2.0
        L[i, i - 1 if i > 0 else 0] += 1.0
        if i < n - 1:
            L[i, i + 1] += 1.0
        else:
            L[i, i] -= 1.0
    L /= d * d
    i = np.arange(n)
    for m in range(n):
        e = np.cos(np.pi * (2 * i + 1) * (2 * m + 1) / (4 * n))
        assert np.allclose(L @ e, k[m] * e, atol=1e-12 * max(1.0, abs(k[m])))
    assert np.all(k < 0)  # no zero mode: the operator is not singular, no pin


def test_io_grid_rules():
    g = io_grid(12, 10, 8)
    assert (g.nx, g.nxg) == (13, 14)
    assert g.shape_u == (13, 12, 11) and g.shape_v == (14, 11, 11) and g.shape_rhs == (13, 10, 10)
    # the nxp cells span Lx (initialization.f90:331-334, Lxp = Lx)
    assert g.x[-1] - g.x[0] == pytest.approx(2 * math.pi, rel=1e-14)
    assert g.dx == pytest.approx(2 * math.pi / 12, rel=1e-14)
    ng = ref_c.channel_grid_arrays(12, 10, 8, g.Lx_i, g.Ly_i, g.Lz_i, g.alpha, 2, 1)
    for name in ("x", "y", "z", "xg", "yg", "zg"):
        assert np.array_equal(ng[name], getattr(g, name)), name
    assert ng["dx"] == g.dx and ng["dz"] == g.dz


@pytest.mark.parametrize("dims", [(8, 5, 6), (6, 3, 9), (12, 1, 4), (4, 6, 1)])
def test_io_solve_matches_dense_operator(dims):
    nxp, nym, nzp = dims
    g = io_grid(*dims)
    pr = R.Projection(g)
    rng = np.random.default_rng(3)
    _, _, _, _, rhs = R.new_fields(g)
    r = rng.standard_normal((nxp, nym, nzp))
    rhs[:nxp, :nym, :nzp] = r
    pr.solve_poisson_equation(rhs)
    A = R.dense_inflow_outflow_operator(g)
    x = np.linalg.solve(A, r.reshape(-1, order="F")).reshape((nxp, nym, nzp), order="F")
    assert np.max(np.abs(rhs[:nxp, :nym, :nzp] - x)) <= 1e-12 * np.max(np.abs(x))
    # periodic z duplicate only (projection.f90:481); no x duplicate (:446-448)
    assert np.array_equal(rhs[: g.nxg - 2, :, g.nzp], rhs[: g.nxg - 2, :, 0])


@pytest.mark.parametrize("oracle", ["numpy", "c"])
def test_uniform_backflow_face_value_gives_a_constant_phi(oracle):
    g = io_grid(8, 6, 4)
    U, V, W, P, rhs = R.new_fields(g)  # u* = 0: no divergence, |u|^2 = 0 at the outlet
    dt, c, p0 = 1e-3, 0.75, 2.5
    Pb = np.full((g.nyg - 2, g.nzg - 2), p0)
    if oracle == "numpy":
        pr = R.Projection(g)
        pr.set_outflow(1.0, 1.0, Pb)
        pr.compute_projection_step(U, V, W, rhs, P, dt, c)
        phib = pr.phib
    else:
        pr = ref_c.CProjection(g)
        pr.set_outflow(1.0, 1.0, Pb)
        pr.compute_projection_step(U, V, W, rhs, P, dt, c)
        phib = pr.phib
    want = -dt * c * p0
    assert np.allclose(phib, want, rtol=1e-15, atol=0)
    assert np.allclose(rhs[: g.nxp, :, : g.nzp], want, rtol=1e-12, atol=0)
    assert max(np.max(np.abs(U)), np.max(np.abs(V)), np.max(np.abs(W))) <= 1e-12
    assert np.allclose(P, want / (dt * c), rtol=1e-12, atol=0)


def _fill_z(g, U, V, W):
    for F in (U, V):
        F[:, :, 0] = F[:, :, g.nzg - 3]; F[:, :, g.nzg - 2] = F[:, :, 1]; F[:, :, g.nzg - 1] = F[:, :, 2]
    W[:, :, 0] = W[:, :, g.nz - 2]; W[:, :, g.nz - 1] = W[:, :, 1]


def io_manufactured(g, seed):
    """u* = curl_h(A) + grad_h(phi_true) with the inflow/outflow x faces: the
    inflow face U(1) takes no gradient (it is never corrected), the outflow
    face U(nx) the gradient toward a zero face value, -2 phi/dx."""
    rng = np.random.default_rng(seed)
    nxp, nym, nzp = g.nxp, g.nym, g.nzp
    y, yg, z, zg, dx = g.y, g.yg, g.z, g.zg, g.dx
    # Az at (x face f = 0..nxp, y face, cell k); Ax at (cell i, y face, z face k+1/2); zero on the walls
    Az = rng.standard_normal((nxp + 1, nym + 1, nzp)); Az[:, 0] = 0; Az[:, nym] = 0
    Ax = rng.standard_normal((nxp, nym + 1, nzp)); Ax[:, 0] = 0; Ax[:, nym] = 0
    dyc = (y[1:] - y[:-1])[None, :, None]
    dzc = (z[1:nzp + 1] - z[0:nzp])[None, None, :]
    u = (Az[:, 1:, :] - Az[:, :-1, :]) / dyc                     # x faces 0..nxp
    v = -(Az[1:] - Az[:-1]) / dx + (Ax - np.roll(Ax, 1, axis=2)) / dzc
    w = -(Ax[:, 1:, :] - Ax[:, :-1, :]) / dyc
    phi = rng.standard_normal((nxp, nym, nzp))
    gu = np.zeros((nxp + 1, nym, nzp))
    gu[1:nxp] = (phi[1:] - phi[:-1]) / dx
    gu[nxp] = -2.0 * phi[nxp - 1] / dx
    dyg = (yg[2:nym + 1] - yg[1:nym])[None, :, None]
    gv = np.zeros_like(v); gv[:, 1:nym, :] = (phi[:, 1:, :] - phi[:, :-1, :]) / dyg
    dzg = (zg[2:nzp + 2] - zg[1:nzp + 1])[None, None, :]
    gw = (np.roll(phi, -1, axis=2) - phi) / dzg
    U, V, W, P, rhs = R.new_fields(g)
    U[0:nxp + 1, 1:nym + 1, 1:nzp + 1] = u + gu
    V[1:nxp + 1, 0:nym + 1, 1:nzp + 1] = v + gv
    W[1:nxp + 1, 1:nym + 1, 1:nzp + 1] = w + gw
    _fill_z(g, U, V, W)
    return U, V, W, rhs, phi, (u, v, w)


# (nym = 1 is left out: the reference's one-cell Dyy, a+b, is not the MAC
# Laplacian between two Neumann walls, so no manufactured pair exists there;
# the dense-operator test covers it)
@pytest.mark.parametrize("dims,alpha", [((8, 6, 8), 1.5), ((12, 9, 6), 2.2), ((4, 3, 1), 1.0)])
def test_io_manufactured_solution_and_idempotence(dims, alpha):
    g = io_grid(*dims, alpha)
    U, V, W, rhs, phi, (u, v, w) = io_manufactured(g, 7)
    pr = R.Projection(g)
    pr.compute_projection_step(U, V, W, rhs)
    nxp, nym, nzp = dims
    assert np.max(np.abs(rhs[:nxp, :nym, :nzp] - phi)) <= 1e-11 * np.max(np.abs(phi))
    assert np.max(np.abs(U[0:nxp + 1, 1:nym + 1, 1:nzp + 1] - u)) <= 1e-10 * np.max(np.abs(u))
    assert np.max(np.abs(V[1:nxp + 1, 0:nym + 1, 1:nzp + 1] - v)) <= 1e-10 * max(np.max(np.abs(v)), 1.0)
    assert np.max(np.abs(W[1:nxp + 1, 1:nym + 1, 1:nzp + 1] - w)) <= 1e-10 * np.max(np.abs(w))
    _fill_z(g, U, V, W)
    assert pr.check_divergence(U, V, W) < 1e-11 * max(np.abs(U).max(), np.abs(V).max()) / g.dx
    U2, V2, W2 = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    pr.compute_projection_step(U2, V2, W2, rhs)
    assert np.max(np.abs(rhs[:nxp, :nym, :nzp])) < 1e-11 * np.max(np.abs(phi))
    assert np.max(np.abs(U2 - U)) < 1e-11 * np.abs(U).max()


def _random_io_fields(g, seed):
    rng = np.random.default_rng(seed)
    U, V, W, P, rhs = R.new_fields(g)
    for A in (U, V, W):
        A[...] = rng.standard_normal(A.shape)
    U += 1.0  # mostly outflow, some backflow cells after the noise
    V[:, 0, :] = 0.0
    V[:, -1, :] = 0.0
    _fill_z(g, U, V, W)
    Pb = rng.standard_normal((g.nyg - 2, g.nzg - 2))
    return U, V, W, P, rhs, Pb


@pytest.mark.parametrize("beta", [0.0, 0.8])
def test_io_projection_is_divergence_free_with_and_without_backflow(beta):
    g = io_grid(12, 8, 6)
    U, V, W, P, rhs, Pb = _random_io_fields(g, 5)
    pr = R.Projection(g)
    pr.set_outflow(beta, 0.9, Pb)
    pr.compute_projection_step(U, V, W, rhs, P, 1e-3, 0.75)
    _fill_z(g, U, V, W)
    scale = max(np.abs(U).max(), np.abs(V).max(), np.abs(W).max()) / g.dx
    assert pr.check_divergence(U, V, W) < 1e-12 * scale * 100


def test_io_pressure_ghosts():
    g = io_grid(6, 5, 4)
    pr = R.Projection(g)
    rng = np.random.default_rng(4)
    _, _, _, P, rhs = R.new_fields(g)
    rhs[: g.nxp, : g.nym, : g.nzp] = rng.standard_normal((g.nxp, g.nym, g.nzp))
    pr.solve_poisson_equation(rhs, P, dt=1e-3, rk_coef33=0.75)
    phi = rhs[: g.nxp, : g.nym, : g.nzp]
    for i in range(1, g.nxg + 1):
        for j in range(1, g.nyg + 1):
            for k in range(1, g.nzg + 1):
                ii = min(max(i - 2, 0), g.nxp - 1)   # Neumann copies at both x ends
                jj = min(max(j - 2, 0), g.nym - 1)
                kk = (k - 2) % g.nzp
                assert P[i - 1, j - 1, k - 1] == phi[ii, jj, kk] / (1e-3 * 0.75)


@pytest.mark.parametrize("dims", [(8, 5, 6), (12, 10, 8), (24, 9, 18), (6, 1, 4), (16, 4, 1)])
@pytest.mark.parametrize("beta", [0.0, 0.6])
def test_c_oracle_matches_numpy_on_inflow_outflow(dims, beta):
    g = io_grid(*dims)
    U, V, W, P, rhs, Pb = _random_io_fields(g, 11)
    a = [x.copy(order="F") for x in (U, V, W, P, rhs)]
    b = [x.copy(order="F") for x in (U, V, W, P, rhs)]
    pn = R.Projection(g)
    pn.set_outflow(beta, 1.2, Pb)
    pn.compute_projection_step(a[0], a[1], a[2], a[4], a[3], 1e-3, 0.75)
    pc = ref_c.CProjection(g)
    pc.set_outflow(beta, 1.2, Pb)
    pc.compute_projection_step(b[0], b[1], b[2], b[4], b[3], 1e-3, 0.75)
    for lab, x, y in zip(("U", "V", "W", "P", "phi"), b, a):
        assert np.max(np.abs(x - y)) <= 1e-13 * max(np.max(np.abs(y)), 1e-300), lab
    if beta > 0:
        assert np.max(np.abs(pc.phib - pn.phib)) <= 1e-15 * np.max(np.abs(pn.phib))
    pc.close()
