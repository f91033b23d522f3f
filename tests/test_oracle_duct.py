"""Pins the 4-wall duct branch of the oracles (x periodic, y walls, z walls:
src/projection.f90:172-218, src/initialization.f90:544-618) by what the
mathematics fixes:
  * the z eigenbasis: v_m = Qz[:, m] / sqrt_w are eigenvectors of the
    assembled z-wall operator Dzz with eigenvalue lambda_z(m); Qz is
    orthonormal; lambda_z is ascending, <= 0, with exactly one zero (the
    constant pressure of the Neumann walls)
  * the solve: phi satisfies L phi + pin(phi) = rhs with L the 7-point
    operator (periodic x with spacing dx, Dyy, Dzz) assembled densely and the
    pin of the kx = 0, lambda_z = 0 mode written in physical space: row 2
    gains 0.5 Dyy(2,2) times the cell-width-weighted mean of phi over x and z
    on that row
  * a divergence-free projection (x ghosts refreshed), including a
    tanh-stretched z grid
  * the C restatement (its own Jacobi eigensolver) equals the numpy one
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import ref_numpy as R

DUCTS = [(8, 6, 5, 1.5, 0.0), (6, 7, 8, 1.2, 1.8), (4, 5, 3, 0.0, 0.0), (12, 4, 6, 2.0, 2.5)]


@pytest.mark.parametrize("dims", DUCTS)
def test_z_eigenbasis(dims):
    g = R.duct_grid(*dims)
    lam, Q, sw = R.z_eigenbasis(g.z, g.zg)
    Dzz = R.dyy_matrix(g.z, g.zg)
    n = Dzz.shape[0]
    assert np.allclose(Q.T @ Q, np.eye(n), atol=1e-13)
    scale = np.max(np.abs(lam))
    for m in range(n):
        v = Q[:, m] / sw
        assert np.max(np.abs(Dzz @ v - lam[m] * v)) <= 1e-12 * scale * np.max(np.abs(v))
    assert np.all(np.diff(lam) >= 0) and np.all(lam <= 1e-12 * scale)
    assert np.sum(np.abs(lam) < 1e-8 * scale) == 1


def _dense_duct(g):
    nxp, nym, nzm = g.nxp, g.nym, g.nzp
    Dyy = R.dyy_matrix(g.y, g.yg)
    Dzz = R.dyy_matrix(g.z, g.zg)
    n = nxp * nym * nzm
    idx = lambda i, j, k: (i % nxp) + nxp * (j + nym * k)
    A = np.zeros((n, n))
    w = g.z[1:nzm + 1] - g.z[:nzm]
    for k in range(nzm):
        for j in range(nym):
            for i in range(nxp):
                r = idx(i, j, k)
                A[r, idx(i + 1, j, k)] += 1.0 / g.dx ** 2
                A[r, idx(i - 1, j, k)] += 1.0 / g.dx ** 2
                A[r, r] -= 2.0 / g.dx ** 2
                for jj in range(nym):
                    A[r, idx(i, jj, k)] += Dyy[j, jj]
                for kk in range(nzm):
                    A[r, idx(i, j, kk)] += Dzz[k, kk]
    # the pin, in physical space: row 2 (j = 0 here) gains 0.5 Dyy(2,2) times
    # the weighted mean over x and z of phi on that row
    for k in range(nzm):
        for i in range(nxp):
            r = idx(i, 0, k)
            for kk in range(nzm):
                for ii in range(nxp):
                    A[r, idx(ii, 0, kk)] += 0.5 * Dyy[0, 0] * w[kk] / (np.sum(w) * nxp)
    return A


@pytest.mark.parametrize("dims", DUCTS)
def test_duct_solve_matches_dense_operator(dims):
    g = R.duct_grid(*dims)
    rng = np.random.default_rng(sum(dims[:3]))
    _, _, _, _, rhs = R.new_fields(g)
    r = rng.standard_normal((g.nxp, g.nym, g.nzp))
    rhs[:g.nxp, :, :g.nzp] = r
    R.Projection(g).solve_poisson_equation(rhs)
    phi = rhs[:g.nxp, :, :g.nzp]
    A = _dense_duct(g)
    lhs = (A @ phi.reshape(-1, order="F")).reshape(phi.shape, order="F")
    assert np.max(np.abs(lhs - r)) <= 1e-10 * np.max(np.abs(r))
    # no z fill: the z ghost position of rhs_p is left as it was (zero)
    assert np.all(rhs[:, :, g.nzp] == 0.0)
    # x duplicate
    assert np.array_equal(rhs[g.nxp], rhs[0])


def _refresh_x(g, U, V, W):
    U[0] = U[g.nx - 2]; U[g.nx - 1] = U[1]
    for F in (V, W):
        F[0] = F[g.nxg - 3]; F[g.nxg - 2] = F[1]; F[g.nxg - 1] = F[2]


@pytest.mark.parametrize("dims", DUCTS)
def test_duct_projection_is_divergence_free(dims):
    g = R.duct_grid(*dims)
    rng = np.random.default_rng(3 + dims[0])
    U, V, W, P, rhs = R.new_fields(g)
    for F in (U, V, W):
        F[...] = rng.standard_normal(F.shape)
    # no flux through the walls (V at the y walls, W at the z walls), as the
    # boundary conditions leave u*: the Neumann problem is then solvable
    V[:, 0, :] = 0.0; V[:, -1, :] = 0.0
    W[:, :, 0] = 0.0; W[:, :, -1] = 0.0
    _refresh_x(g, U, V, W)
    pr = R.Projection(g)
    pr.compute_projection_step(U, V, W, rhs, P, 1e-3, 0.75)
    _refresh_x(g, U, V, W)
    assert pr.check_divergence(U, V, W) <= 1e-10
    # P: Neumann ghosts at the z walls
    assert np.array_equal(P[:, :, 0], P[:, :, 1]) and np.array_equal(P[:, :, -1], P[:, :, -2])


@pytest.mark.parametrize("dims", DUCTS)
def test_c_oracle_equals_numpy_duct(dims):
    from oracle import ref_c

    g = R.duct_grid(*dims)
    rng = np.random.default_rng(11 + dims[1])
    U, V, W, P, rhs = R.new_fields(g)
    for F in (U, V, W):
        F[...] = rng.standard_normal(F.shape)
    a = [x.copy(order="F") for x in (U, V, W, P, rhs)]
    b = [x.copy(order="F") for x in (U, V, W, P, rhs)]
    R.Projection(g).compute_projection_step(a[0], a[1], a[2], a[4], a[3], 1e-3, 0.75)
    op = ref_c.CProjection(g)
    op.compute_projection_step(b[0], b[1], b[2], b[4], b[3], 1e-3, 0.75)
    op.close()
    for x, y in zip(a, b):
        assert np.max(np.abs(x - y)) <= 1e-12 * max(1.0, np.max(np.abs(y)))


def test_duct_boundary_conditions():
    # boundary_conditions.f90:186-260 with z walls (:215-223, :737-758): W is
    # zero on the wall faces, U and V vanish at the wall (ghost mirrored), x
    # wraps hold, y Robin rows as in the channel; C == numpy bit for bit
    from oracle import ref_c

    g = R.duct_grid(8, 6, 5, 1.5, 1.0)
    rng = np.random.default_rng(2)
    U, V, W = (np.asfortranarray(rng.standard_normal(s)) for s in (g.shape_u, g.shape_v, g.shape_w))
    al = (np.asfortranarray(rng.uniform(0, 0.2, (g.nx, 2, g.nzg))),
          np.asfortranarray(rng.uniform(0, 0.2, (g.nxg, 2, g.nzg))),
          np.asfortranarray(rng.uniform(0, 0.2, (g.nxg, 2, g.nz))))
    a = [x.copy(order="F") for x in (U, V, W)]
    R.apply_boundary_conditions(g, *a, *al)
    Ua, Va, Wa = a
    assert np.all(Wa[:, :, 0] == 0.0) and np.all(Wa[:, :, -1] == 0.0)
    for F in (Ua, Va):
        assert np.array_equal(F[:, 1:-1, 0], -F[:, 1:-1, 1]) and np.array_equal(F[:, 1:-1, -1], -F[:, 1:-1, -2])
    assert np.array_equal(Ua[0], Ua[g.nx - 2]) and np.array_equal(Ua[g.nx - 1], Ua[1])
    b = [x.copy(order="F") for x in (U, V, W)]
    op = ref_c.CProjection(g)
    op.apply_boundary_conditions(*b, *al)
    op.close()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


# spanwise walls with a periodic y (x periodic, y periodic, z walls)
ZWPY = [(8, 5, 6, 0.0), (6, 9, 7, 1.5), (4, 4, 3, 0.0), (10, 7, 5, 2.0)]


def _dense_zwall_py(g):
    nxp, nyp, nzm = g.nxp, g.nym - 1, g.nzp
    dy = float(g.y[1] - g.y[0])
    Dzz = R.dyy_matrix(g.z, g.zg)
    n = nxp * nyp * nzm
    idx = lambda i, j, k: (i % nxp) + nxp * ((j % nyp) + nyp * k)
    A = np.zeros((n, n))
    for k in range(nzm):
        for j in range(nyp):
            for i in range(nxp):
                r = idx(i, j, k)
                for (a, b), h in (((1, 0), g.dx), ((0, 1), dy)):
                    A[r, idx(i + a, j + b, k)] += 1.0 / h ** 2
                    A[r, idx(i - a, j - b, k)] += 1.0 / h ** 2
                    A[r, r] -= 2.0 / h ** 2
                for kk in range(nzm):
                    A[r, idx(i, j, kk)] += Dzz[k, kk]
    # the pin of the kx = ky = 0 mode on its first z row: 0.5 Dzz(2,2) times
    # the x-y mean of phi on the first z row
    for j in range(nyp):
        for i in range(nxp):
            for jj in range(nyp):
                for ii in range(nxp):
                    A[idx(i, j, 0), idx(ii, jj, 0)] += 0.5 * Dzz[0, 0] / (nxp * nyp)
    return A


@pytest.mark.parametrize("dims", ZWPY)
def test_zwall_periodic_y_solve_matches_dense_operator(dims):
    g = R.zwall_periodic_y_grid(*dims)
    rng = np.random.default_rng(sum(dims[:3]))
    _, _, _, _, rhs = R.new_fields(g)
    r = rng.standard_normal((g.nxp, g.nym, g.nzp))
    rhs[:g.nxp, :, :g.nzp] = r
    R.Projection(g).solve_poisson_equation(rhs)
    nyp = g.nym - 1
    phi = rhs[:g.nxp, :nyp, :g.nzp]
    lhs = (_dense_zwall_py(g) @ phi.reshape(-1, order="F")).reshape(phi.shape, order="F")
    assert np.max(np.abs(lhs - r[:, :nyp, :])) <= 1e-10 * np.max(np.abs(r))
    # the duplicate row, the x duplicate, no z fill
    assert np.array_equal(rhs[:, g.nym - 1, :], rhs[:, 0, :]) and np.array_equal(rhs[g.nxp], rhs[0])
    assert np.all(rhs[:g.nxp, :, g.nzp] == 0.0)


def _refresh_xy(g, U, V, W):
    _refresh_x(g, U, V, W)
    V[:, 0, :] = V[:, g.ny - 2, :]; V[:, g.ny - 1, :] = V[:, 1, :]
    for F in (U, W):
        F[:, 0, :] = F[:, g.nyg - 3, :]; F[:, g.nyg - 2, :] = F[:, 1, :]; F[:, g.nyg - 1, :] = F[:, 2, :]


@pytest.mark.parametrize("dims", ZWPY)
def test_zwall_periodic_y_projection_is_divergence_free_and_c_equals_numpy(dims):
    from oracle import ref_c

    g = R.zwall_periodic_y_grid(*dims)
    rng = np.random.default_rng(40 + dims[0])
    U, V, W, P, rhs = R.new_fields(g)
    for F in (U, V, W):
        F[...] = rng.standard_normal(F.shape)
    W[:, :, 0] = 0.0; W[:, :, -1] = 0.0
    _refresh_xy(g, U, V, W)
    a = [x.copy(order="F") for x in (U, V, W, P, rhs)]
    b = [x.copy(order="F") for x in (U, V, W, P, rhs)]
    pr = R.Projection(g)
    pr.compute_projection_step(a[0], a[1], a[2], a[4], a[3], 1e-3, 0.75)
    op = ref_c.CProjection(g)
    op.compute_projection_step(b[0], b[1], b[2], b[4], b[3], 1e-3, 0.75)
    op.close()
    for x, y in zip(a, b):
        assert np.max(np.abs(x - y)) <= 1e-12 * max(1.0, np.max(np.abs(y)))
    _refresh_xy(g, a[0], a[1], a[2])
    assert pr.check_divergence(a[0], a[1], a[2]) <= 1e-10
