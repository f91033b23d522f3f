"""Pins the periodic-y branch of the oracles (y_bc_type = 0 with periodic x
and z, a uniform y grid: src/initialization.f90:467-491, projection.f90:
246-273, :350-354) by what the mathematics fixes:
  * kyy: eigenvalues of the periodic 3-point Laplacian on nyp = nym-1 rows
  * the solve: with the (0,0,0) mode set to zero, phi on the independent
    cells is the zero-mean solution of L phi = r - mean(r), L the triply
    periodic 7-point operator (spacings dx, dy, dz) assembled densely
  * the duplicate row (and the x and z duplicates) copy their sources
  * a divergence-free result: after the projection the divergence, with the
    periodic ghosts refreshed in x, y and z, is at round-off
  * the C restatement equals the numpy restatement (1e-13)
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import ref_c
from oracle import ref_numpy as R


def _dense_periodic(g):
    nxp, nyp, nzp = g.nxp, g.nym - 1, g.nzp
    n = nxp * nyp * nzp
    A = np.zeros((n, n))
    dy = float(g.y[1] - g.y[0])

    def idx(i, j, k):
        return (i % nxp) + nxp * ((j % nyp) + nyp * (k % nzp))

    for k in range(nzp):
        for j in range(nyp):
            for i in range(nxp):
                r = idx(i, j, k)
                for (a, b, c), h in (((1, 0, 0), g.dx), ((0, 1, 0), dy), ((0, 0, 1), g.dz)):
                    A[r, idx(i + a, j + b, k + c)] += 1.0 / h ** 2
                    A[r, idx(i - a, j - b, k - c)] += 1.0 / h ** 2
                    A[r, r] -= 2.0 / h ** 2
    return A


@pytest.mark.parametrize("n,d", [(8, 0.25), (9, 0.1), (6, 1.3)])
def test_kyy_eigen_identity(n, d):
    k = R.wavenumbers(n, d)
    L = np.zeros((n, n))
    for i in range(n):
        L[i, i] = -2.0
        L[i, (i + 1) % n] += 1.0
        L[i, (i - 1) % n] += 1.0
    L /= d * d
    j = np.arange(n)
    for m in range(n):
        e = np.exp(2j * np.pi * j * m / n)
        assert np.allclose(L @ e, k[m] * e, atol=1e-12 * max(1.0, abs(k[m])))


@pytest.mark.parametrize("dims", [(8, 7, 6), (6, 10, 4), (4, 4, 9), (12, 3, 8)])
@pytest.mark.parametrize("oracle", ["numpy", "c"])
def test_periodic_y_solve_matches_dense_operator(dims, oracle):
    g = R.periodic_y_grid(*dims)
    nxp, nym, nzp = dims
    nyp = nym - 1
    rng = np.random.default_rng(sum(dims))
    _, _, _, _, rhs = R.new_fields(g)
    r = rng.standard_normal((nxp, nym, nzp))
    rhs[:nxp, :nym, :nzp] = r
    if oracle == "numpy":
        R.Projection(g).solve_poisson_equation(rhs)
    else:
        op = ref_c.CProjection(g)
        op.solve_poisson_equation(rhs)
        op.close()
    phi = rhs[:nxp, :nyp, :nzp]
    ri = r[:, :nyp, :]
    A = _dense_periodic(g)
    lhs = (A @ phi.reshape(-1, order="F")).reshape(phi.shape, order="F")
    scale = np.max(np.abs(ri))
    assert np.max(np.abs(lhs - (ri - ri.mean()))) <= 1e-11 * scale
    assert abs(phi.mean()) <= 1e-13 * np.max(np.abs(phi))
    # duplicates: row nym-1 (:350-354, whole x-z extent), x column nxp, z plane nzp
    assert np.array_equal(rhs[:, nym - 1, :], rhs[:, 0, :])
    assert np.array_equal(rhs[g.nxp, :, :], rhs[0, :, :])
    assert np.array_equal(rhs[: g.nxg - 2, :, g.nzp], rhs[: g.nxg - 2, :, 0])


def _refresh_periodic(g, U, V, W):
    # x and z: boundary_conditions.f90:275-284, :591-600; y: :762-781
    U[0] = U[g.nx - 2]; U[g.nx - 1] = U[1]
    for F in (V, W):
        F[0] = F[g.nxg - 3]; F[g.nxg - 2] = F[1]; F[g.nxg - 1] = F[2]
    for F in (U, V):
        F[:, :, 0] = F[:, :, g.nzg - 3]; F[:, :, g.nzg - 2] = F[:, :, 1]; F[:, :, g.nzg - 1] = F[:, :, 2]
    W[:, :, 0] = W[:, :, g.nz - 2]; W[:, :, g.nz - 1] = W[:, :, 1]
    V[:, 0, :] = V[:, g.ny - 2, :]; V[:, g.ny - 1, :] = V[:, 1, :]
    for F in (U, W):
        F[:, 0, :] = F[:, g.nyg - 3, :]; F[:, g.nyg - 2, :] = F[:, 1, :]; F[:, g.nyg - 1, :] = F[:, 2, :]


@pytest.mark.parametrize("oracle", ["numpy", "c"])
def test_periodic_y_projection_is_divergence_free(oracle):
    g = R.periodic_y_grid(8, 9, 6)
    rng = np.random.default_rng(4)
    U, V, W, P, rhs = R.new_fields(g)
    for F in (U, V, W):
        F[...] = rng.standard_normal(F.shape)
    _refresh_periodic(g, U, V, W)
    if oracle == "numpy":
        pr = R.Projection(g)
        pr.compute_projection_step(U, V, W, rhs, P, 1e-3, 0.75)
    else:
        pr = ref_c.CProjection(g)
        pr.compute_projection_step(U, V, W, rhs, P, 1e-3, 0.75)
    _refresh_periodic(g, U, V, W)
    md = pr.check_divergence(U, V, W)
    if oracle == "c":
        pr.close()
    assert md <= 1e-11, md


def test_c_oracle_equals_numpy_periodic_y():
    g = R.periodic_y_grid(12, 10, 8)
    rng = np.random.default_rng(7)
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
        assert np.max(np.abs(x - y)) <= 1e-13 * max(1.0, np.max(np.abs(y)))


def test_periodic_y_boundary_conditions():
    # boundary_conditions.f90:186-260 with y_bc_type = 0 (:229-233): every y
    # ghost row equals its periodic source row; the interior is untouched
    g = R.periodic_y_grid(8, 7, 6)
    rng = np.random.default_rng(21)
    U, V, W = (np.asfortranarray(rng.standard_normal(s)) for s in (g.shape_u, g.shape_v, g.shape_w))
    U0, V0, W0 = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    a = [x.copy(order="F") for x in (U, V, W)]
    R.apply_boundary_conditions(g, *a, None, None, None)
    Ua, Va, Wa = a
    ny, nyg = g.ny, g.nyg
    assert np.array_equal(Va[:, 0, :], Va[:, ny - 2, :]) and np.array_equal(Va[:, ny - 1, :], Va[:, 1, :])
    for F in (Ua, Wa):
        assert np.array_equal(F[:, 0, :], F[:, nyg - 3, :])
        assert np.array_equal(F[:, nyg - 2, :], F[:, 1, :]) and np.array_equal(F[:, nyg - 1, :], F[:, 2, :])
    assert np.array_equal(Ua[1:-1, 1:-2, 1:-2], U0[1:-1, 1:-2, 1:-2])
    assert np.array_equal(Va[1:-2, 1:-1, 1:-2], V0[1:-2, 1:-1, 1:-2])
    assert np.array_equal(Wa[1:-2, 1:-2, 1:-1], W0[1:-2, 1:-2, 1:-1])
    # x and z wraps still hold after the y wrap (the second pass, :247-258)
    assert np.array_equal(Ua[0], Ua[g.nx - 2]) and np.array_equal(Wa[:, :, 0], Wa[:, :, g.nz - 2])
    b = [x.copy(order="F") for x in (U, V, W)]
    op = ref_c.CProjection(g)
    op.apply_boundary_conditions(*b, None, None, None)
    op.close()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
