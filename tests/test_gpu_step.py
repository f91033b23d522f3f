"""GPU parity of the step either side of the projection (channel, one rank):
the RK3 stage update (time_integration.f90:256-258, :394-396, :526-531),
apply_boundary_conditions (boundary_conditions.f90:186-260) and one whole
substage around the projection (time_integration.f90:256-304), against the C
oracle on identical device-resident inputs.

Bars: the update and the boundary conditions bit-identical (same operations);
the substage (which contains the projection) U, V, W, phi and P to relative
L-inf 1e-12.
"""
from __future__ import annotations

import numpy as np
import pytest

import dfp
from oracle import ref_c
from _cases import make_grid, rel_linf

pytestmark = pytest.mark.gpu

DT = 1e-3
# initialization.f90:655-660
RK_COEF = np.array([[8 / 15, 0, 0], [1 / 4, 5 / 12, 0], [1 / 4, 0, 3 / 4]])
GRIDS = [(16, 10, 8, 1.5), (12, 1, 1, 1.0), (24, 9, 18, 2.0)]


def _interior_shapes(g):
    return ((g.nx - 2, g.nyg - 2, g.nzg - 2), (g.nxg - 2, g.ny - 2, g.nzg - 2), (g.nxg - 2, g.nyg - 2, g.nz - 2))


def _inputs(g, seed):
    rng = np.random.default_rng(seed)
    U, V, W = dfp.synth_fields(g, seed, 0.2)
    Uo, Vo, Wo = dfp.synth_fields(g, seed + 1, 0.2)
    F = [[np.asfortranarray(rng.standard_normal(sh)) for _ in range(3)] for sh in _interior_shapes(g)]
    al = (np.asfortranarray(rng.uniform(0, 0.2, (g.nx, 2, g.nzg))),
          np.asfortranarray(rng.uniform(0, 0.2, (g.nxg, 2, g.nzg))),
          np.asfortranarray(rng.uniform(0, 0.2, (g.nxg, 2, g.nz))))
    return [U, V, W], [Uo, Vo, Wo], F, al


def _dev(a):
    return dfp.DeviceArray.from_host(a)


@pytest.mark.parametrize("dims", GRIDS, ids=lambda d: f"{d[0]}x{d[1]}x{d[2]}")
@pytest.mark.parametrize("stage", [1, 2, 3])
def test_rk_update_and_boundary_conditions_bit_identical(dims, stage):
    g = make_grid(*dims)
    print(f"step {dims} stage {stage}", flush=True)
    F3, Fo, F, al = _inputs(g, 10 * stage + dims[0])
    op = ref_c.CProjection(g)
    o = [a.copy(order="F") for a in F3]
    op.rk_update(stage, *o, *Fo, *F, RK_COEF[stage - 1], DT)
    plan = dfp.Plan(g)
    d = [_dev(a) for a in F3]
    plan.rk_update(stage, *d, *[_dev(a) for a in Fo], *[[_dev(x) for x in f] for f in F], RK_COEF[stage - 1], DT)
    for x, y in zip(d, o):
        assert np.array_equal(x.download(), y)
    op.apply_boundary_conditions(*o, *al)
    plan.apply_bc(*d, *[_dev(a) for a in al])
    for x, y in zip(d, o):
        assert np.array_equal(x.download(), y)
    op.close()
    plan.close()


@pytest.mark.parametrize("stage", [1, 2, 3])
def test_substep_matches_the_oracle_sequence(stage):
    g = make_grid(32, 24, 16, 1.8)
    F3, Fo, F, al = _inputs(g, 700 + stage)
    P = np.zeros(g.shape_p, order="F")
    rhs = np.zeros(g.shape_rhs, order="F")
    save_p = stage == 3
    c = RK_COEF[stage - 1][stage - 1]
    op = ref_c.CProjection(g)
    o = [a.copy(order="F") for a in (*F3, P, rhs)]
    op.rk_update(stage, o[0], o[1], o[2], *Fo, *F, RK_COEF[stage - 1], DT)
    op.apply_boundary_conditions(o[0], o[1], o[2], *al)
    op.compute_projection_step(o[0], o[1], o[2], o[4], o[3] if save_p else None, DT, c)
    op.apply_boundary_conditions(o[0], o[1], o[2], *al)
    op.close()
    plan = dfp.Plan(g)
    d = [_dev(a) for a in (*F3, P, rhs)]
    plan.substep(stage, d[0], d[1], d[2], *[_dev(a) for a in Fo], *[[_dev(x) for x in f] for f in F],
                 RK_COEF[stage - 1], *[_dev(a) for a in al], d[4], d[3], DT, save_p=save_p)
    for lab, x, y in zip(("U", "V", "W", "P", "phi"), d, o):
        err = rel_linf(x.download(), y)
        assert err <= 1e-12, (stage, lab, err)
    plan.close()


def test_step_entry_points_refuse_host_fields_and_inflow_plans():
    g = make_grid(8, 6, 4, 1.5)
    plan = dfp.Plan(g)
    F3, Fo, F, al = _inputs(g, 5)
    with pytest.raises(ValueError):
        plan.rk_update(1, *F3, *Fo, *F, RK_COEF[0], DT)
    with pytest.raises(dfp.DfpError) as ei:
        plan.apply_bc(*F3, *[_dev(a) for a in al])
    assert ei.value.code == -1
    plan.close()
    gi = dfp.channel_grid(8, 6, 4, 1.5, x_bc_type=1)
    pi = dfp.Plan(gi)
    U, V, W = (dfp.DeviceArray(s, zero=True) for s in (gi.shape_u, gi.shape_v, gi.shape_w))
    ali = (dfp.DeviceArray((gi.nx, 2, gi.nzg), zero=True), dfp.DeviceArray((gi.nxg, 2, gi.nzg), zero=True),
           dfp.DeviceArray((gi.nxg, 2, gi.nz), zero=True))
    with pytest.raises(dfp.DfpError) as ei:
        pi.apply_bc(U, V, W, *ali)
    assert ei.value.code == -2
    pi.close()
