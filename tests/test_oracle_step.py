"""Pins the oracle's restatement of the step either side of the projection
(channel case, one rank): the RK3 stage update (time_integration.f90:256-258,
:394-396, :526-531) and apply_boundary_conditions (boundary_conditions.f90:
186-260: periodic x and z wraps, Robin slip-length y walls), by what the
mathematics fixes:
  * the stage update with one right-hand side for every stage is
    Uo + dt*rk_t(s)*F (the reference's rk_coef rows sum to rk_t(s) = 8/15,
    2/3, 1, initialization.f90:655-666), and it is linear in each F_s with
    weight dt*rk_coef(s, s'); ghost planes are untouched
  * the wraps leave every ghost plane equal to its periodic source plane
  * the Robin rows satisfy the slip condition they discretise, u = alpha
    du/dn with n the normal into the fluid: at a centred field (U, W) the
    wall value (F(1)+F(2))/2 equals alpha (F(2)-F(1))/(yg(2)-yg(1)) (and
    alpha (F(nyg-1)-F(nyg))/(yg(nyg)-yg(nyg-1)) at the upper wall); at V,
    V(1) = alpha (V(2)-V(1))/dy;
    alpha = 0 gives no slip (U(1) = -U(2), V(1) = 0) and a large alpha free
    slip (U(1) -> U(2))
  * the C restatement equals the numpy restatement bit for bit
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import ref_c
from oracle import ref_numpy as R

# initialization.f90:655-666
RK_COEF = np.array([[8 / 15, 0, 0], [1 / 4, 5 / 12, 0], [1 / 4, 0, 3 / 4]])
RK_T = (8 / 15, 2 / 3, 1.0)


def _grid():
    return R.channel_grid(12, 7, 8, 1.6)


def _interior_shapes(g):
    return ((g.nx - 2, g.nyg - 2, g.nzg - 2), (g.nxg - 2, g.ny - 2, g.nzg - 2), (g.nxg - 2, g.nyg - 2, g.nz - 2))


def _rand(shape, rng):
    return np.asfortranarray(rng.standard_normal(shape))


@pytest.mark.parametrize("stage", [1, 2, 3])
def test_rk_update_constant_rhs_and_linearity(stage):
    g = _grid()
    rng = np.random.default_rng(stage)
    shapes = (g.shape_u, g.shape_v, g.shape_w)
    Uo = [_rand(s, rng) for s in shapes]
    out = [_rand(s, rng) for s in shapes]
    const = [_rand(s, rng) for s in _interior_shapes(g)]
    dt = 2e-3
    # one right-hand side for every stage
    Fs = [[c] * 3 for c in const]
    got = [a.copy(order="F") for a in out]
    R.rk_update(g, stage, *got, *Uo, *Fs, RK_COEF[stage - 1], dt)
    for a, a0, o, c in zip(got, out, Uo, const):
        inner = (slice(1, -1),) * 3
        np.testing.assert_allclose(a[inner], o[inner] + dt * RK_T[stage - 1] * c, rtol=0, atol=1e-15 * 4)
        ghost = np.ones(a.shape, bool)
        ghost[inner] = False
        assert np.array_equal(a[ghost], a0[ghost])
    # linearity in each right-hand side, weight dt*rk_coef(stage, s)
    base = [[_rand(sh, rng) for _ in range(3)] for sh in _interior_shapes(g)]
    r0 = [a.copy(order="F") for a in out]
    R.rk_update(g, stage, *r0, *Uo, *base, RK_COEF[stage - 1], dt)
    for s in range(stage):
        pert = [[f.copy(order="F") for f in b] for b in base]
        d = [_rand(sh, rng) for sh in _interior_shapes(g)]
        for b, dd in zip(pert, d):
            b[s] = b[s] + dd
        r1 = [a.copy(order="F") for a in out]
        R.rk_update(g, stage, *r1, *Uo, *pert, RK_COEF[stage - 1], dt)
        for a1, a0, dd in zip(r1, r0, d):
            inner = (slice(1, -1),) * 3
            np.testing.assert_allclose(a1[inner] - a0[inner], dt * RK_COEF[stage - 1][s] * dd, rtol=0, atol=1e-14)


def _alphas(g, rng, scale=1.0):
    return (np.asfortranarray(scale * rng.uniform(0.0, 0.3, (g.nx, 2, g.nzg))),
            np.asfortranarray(scale * rng.uniform(0.0, 0.3, (g.nxg, 2, g.nzg))),
            np.asfortranarray(scale * rng.uniform(0.0, 0.3, (g.nxg, 2, g.nz))))


def test_boundary_conditions_wraps_and_slip_condition():
    g = _grid()
    rng = np.random.default_rng(7)
    U, V, W = (_rand(s, rng) for s in (g.shape_u, g.shape_v, g.shape_w))
    ax, ay, az = _alphas(g, rng)
    U0, V0, W0 = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    R.apply_boundary_conditions(g, U, V, W, ax, ay, az)
    # interior untouched (centred periodic directions also hold a duplicate
    # plane n-1, which the wraps fill)
    assert np.array_equal(U[1:-1, 1:-1, 1:-2], U0[1:-1, 1:-1, 1:-2])
    assert np.array_equal(V[1:-2, 1:-1, 1:-2], V0[1:-2, 1:-1, 1:-2])
    assert np.array_equal(W[1:-2, 1:-1, 1:-1], W0[1:-2, 1:-1, 1:-1])
    # every ghost plane equals its periodic source plane (x faces: 1 <- nx-1, nx <- 2;
    # centres: 1 <- n-2, n-1 <- 2, n <- 3; z alike)
    nx, nxg, nz, nzg = g.nx, g.nxg, g.nz, g.nzg
    assert np.array_equal(U[0], U[nx - 2]) and np.array_equal(U[nx - 1], U[1])
    for F in (V, W):
        assert np.array_equal(F[0], F[nxg - 3]) and np.array_equal(F[nxg - 2], F[1]) and np.array_equal(F[nxg - 1], F[2])
    for F in (U, V):
        assert np.array_equal(F[:, :, 0], F[:, :, nzg - 3]) and np.array_equal(F[:, :, nzg - 2], F[:, :, 1])
        assert np.array_equal(F[:, :, nzg - 1], F[:, :, 2])
    assert np.array_equal(W[:, :, 0], W[:, :, nz - 2]) and np.array_equal(W[:, :, nz - 1], W[:, :, 1])
    # slip condition at the independent (x, z) positions (the wraps then copy
    # these rows into the ghost and duplicate planes): u_wall = alpha du/dn,
    # n the wall normal into the fluid (+y at the lower wall, -y at the upper)
    g1, g2 = g.yg[1] - g.yg[0], g.yg[-1] - g.yg[-2]
    d1, d2 = g.y[1] - g.y[0], g.y[-1] - g.y[-2]
    for F, a, sl in ((U, ax, (slice(1, -1), slice(1, -2))), (W, az, (slice(1, -2), slice(1, -1)))):
        lo, hi = F[:, 0, :][sl], F[:, -1, :][sl]
        n1, n2 = F[:, 1, :][sl], F[:, -2, :][sl]
        np.testing.assert_allclose(0.5 * (lo + n1), a[:, 0, :][sl] * (n1 - lo) / g1, rtol=0, atol=1e-13)
        np.testing.assert_allclose(0.5 * (hi + n2), a[:, 1, :][sl] * (n2 - hi) / g2, rtol=0, atol=1e-13)
    sl = (slice(1, -2), slice(1, -2))
    np.testing.assert_allclose(V[:, 0, :][sl], ay[:, 0, :][sl] * (V[:, 1, :][sl] - V[:, 0, :][sl]) / d1, atol=1e-13)
    np.testing.assert_allclose(V[:, -1, :][sl], ay[:, 1, :][sl] * (V[:, -2, :][sl] - V[:, -1, :][sl]) / d2,
                               atol=1e-13)


@pytest.mark.parametrize("limit", ["no_slip", "free_slip"])
def test_boundary_conditions_slip_limits(limit):
    g = _grid()
    rng = np.random.default_rng(9)
    U, V, W = (_rand(s, rng) for s in (g.shape_u, g.shape_v, g.shape_w))
    scale = 0.0 if limit == "no_slip" else 1e14
    ax, ay, az = _alphas(g, rng, scale)
    if limit == "free_slip":
        ax, ay, az = (a + 1e13 for a in (ax, ay, az))
    R.apply_boundary_conditions(g, U, V, W, ax, ay, az)
    sl = (slice(1, -1), slice(1, -1))
    if limit == "no_slip":
        for F in (U, W):
            assert np.array_equal(F[:, 0, :][sl], -F[:, 1, :][sl]) and np.array_equal(F[:, -1, :][sl], -F[:, -2, :][sl])
        assert np.all(V[:, 0, :][sl] == 0.0) and np.all(V[:, -1, :][sl] == 0.0)
    else:
        for F in (U, W):
            np.testing.assert_allclose(F[:, 0, :][sl], F[:, 1, :][sl], rtol=1e-11, atol=1e-11)
        np.testing.assert_allclose(V[:, 0, :][sl], V[:, 1, :][sl], rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("stage", [1, 2, 3])
def test_c_step_oracle_equals_numpy(stage):
    g = _grid()
    rng = np.random.default_rng(100 + stage)
    shapes = (g.shape_u, g.shape_v, g.shape_w)
    Uo = [_rand(s, rng) for s in shapes]
    out = [_rand(s, rng) for s in shapes]
    F = [[_rand(sh, rng) for _ in range(3)] for sh in _interior_shapes(g)]
    a = [x.copy(order="F") for x in out]
    b = [x.copy(order="F") for x in out]
    R.rk_update(g, stage, *a, *Uo, *F, RK_COEF[stage - 1], 1e-3)
    op = ref_c.CProjection(g)
    op.rk_update(stage, *b, *Uo, *F, RK_COEF[stage - 1], 1e-3)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    al = _alphas(g, rng)
    R.apply_boundary_conditions(g, *a, *al)
    op.apply_boundary_conditions(*b, *al)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    op.close()
