"""GPU parity of the x inflow/outflow branch (x_bc_type = 1: DCT-IV in x, no
pin, outlet backflow term, outflow face correction; src/projection.f90:65-96,
:126-129, :333-339, :413-422, :578-589) against the C oracle on identical
inputs.

Bars (written here, as the tests apply them):
  * divergence, correction (without the backflow term), max-divergence:
    bit-identical (same operations)
  * phi (rhs_p), P, phib and the corrected U, V, W: relative L-inf <= 1e-12
  * post-projection max|div u| <= 1e-12 * max|u| / min spacing (after the
    periodic z ghosts of the velocity are refreshed)
"""
from __future__ import annotations

import numpy as np
import pytest

import dfp
from oracle import ref_c
from oracle import ref_numpy as R
from _cases import div_bound, rel_linf

pytestmark = pytest.mark.gpu

TOL = 1e-12
DT = 1e-3
C3 = 3.0 / 4.0

# (name, (nxp, nym, nzp), alpha)
IO_CASES = [
    ("tiny_8x5x6", (8, 5, 6), 1.5),
    ("asym_12x10x8", (12, 10, 8), 2.0),
    ("radix3_24x9x18", (24, 9, 18), 1.2),
    ("odd_nym_32x33x48", (32, 33, 48), 2.5),
    ("nxp2_2x4x4", (2, 4, 4), 1.0),
    ("nzp1_6x7x1", (6, 7, 1), 1.0),
    ("one_cell_8x1x12", (8, 1, 12), 1.5),
    ("radix3x_48x6x12", (48, 6, 12), 1.7),
]


def io_grid(nxp, nym, nzp, alpha):
    return dfp.channel_grid(nxp, nym, nzp, alpha, x_bc_type=1)


def io_fields(g, seed, sigma):
    U, V, W = dfp.synth_fields(g, seed, sigma)
    return U, V, W, np.zeros(g.shape_p, order="F"), np.zeros(g.shape_rhs, order="F")


def backflow_state(g, seed):
    """An outlet face pressure Pb(2:nyg-1, 2:nzg-1) and a plane-mean outlet
    speed; the synthetic outlet velocity has both signs, so the smoothed
    Heaviside of the backflow term is exercised on either side."""
    rng = np.random.default_rng(seed)
    Pb = np.asfortranarray(rng.standard_normal((g.nyg - 2, g.nzg - 2)))
    return Pb, 0.8


def run_both(g, seed, sigma, beta, c=C3, save_p=True):
    U, V, W, P, rhs = io_fields(g, seed, sigma)
    Pb, uc = backflow_state(g, seed + 1)
    o = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    op = ref_c.CProjection(g)
    if beta > 0:
        op.set_outflow(beta, uc, Pb)
    op.compute_projection_step(o[0], o[1], o[2], o[4], o[3] if save_p else None, DT, c)
    phib_o = op.phib if beta > 0 else None
    op.close()
    plan = dfp.Plan(g)
    phib_d = np.zeros((g.nyg - 2, g.nzg - 2), order="F")
    if beta > 0:
        plan.set_outflow(beta, uc, Pb, phib_d)
    d = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    plan.project(d[0], d[1], d[2], d[4], d[3] if save_p else None, DT, c, save_p=save_p)
    plan.close()
    return o, d, phib_o, phib_d


@pytest.mark.parametrize("beta", [0.0, 0.7])
@pytest.mark.parametrize("name,dims,alpha", IO_CASES, ids=[c[0] for c in IO_CASES])
def test_inflow_project_small_cases(name, dims, alpha, beta):
    g = io_grid(*dims, alpha)
    print(f"inflow case {name} beta {beta}", flush=True)
    o, d, fo, fd = run_both(g, 300 + dims[0] + dims[2], 0.1, beta)
    for lab, a, b in zip(("U", "V", "W", "P", "phi"), d, o):
        err = rel_linf(a, b)
        assert err <= TOL, f"{name} beta {beta} {lab}: rel-Linf {err:.3e}"
    if beta > 0:
        assert rel_linf(fd, fo) <= TOL


def test_inflow_stages_bit_identical_and_solve():
    g = io_grid(24, 16, 12, 1.5)
    U, V, W, P, rhs = io_fields(g, 11, 0.1)
    op = ref_c.CProjection(g)
    plan = dfp.Plan(g)
    # divergence (projection.f90:37-63)
    r_o = rhs.copy(order="F"); r_d = rhs.copy(order="F")
    op.compute_pseudo_pressure_rhs(U, V, W, r_o)
    plan.divergence(U, V, W, r_d)
    assert np.array_equal(r_o, r_d)
    # solve (DCT-IV in x, no pin) with P and its Neumann x ghosts
    P_o = P.copy(order="F"); P_d = P.copy(order="F")
    op.solve_poisson_equation(r_o, P_o, DT, C3)
    plan.solve(r_d, P_d, DT, C3, save_p=True)
    assert rel_linf(r_d, r_o) <= TOL
    assert rel_linf(P_d, P_o) <= TOL
    # correction from the same phi, outlet face without the backflow term
    Uo, Vo, Wo = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    Ud, Vd, Wd = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    op.project_velocity(Uo, Vo, Wo, r_o)
    plan.correct(Ud, Vd, Wd, r_o)
    assert np.array_equal(Uo, Ud) and np.array_equal(Vo, Vd) and np.array_equal(Wo, Wd)
    assert plan.max_divergence(U, V, W) == op.check_divergence(U, V, W)
    op.close(); plan.close()


def test_inflow_backflow_stage_and_face_correction():
    """dfp_outflow_backflow against the numpy restatement of
    set_outflow_backflow_pressure (tanh of two libraries: 1e-14), then the
    outlet face correction with that phib, bit for bit."""
    g = io_grid(16, 10, 8, 1.8)
    U, V, W, P, rhs = io_fields(g, 21, 0.3)
    Pb, uc = backflow_state(g, 5)
    beta = 0.9
    c = 5.0 / 12.0
    rp = R.Projection(R.channel_grid(16, 10, 8, 1.8, x_bc_type=1))
    rp.set_outflow(beta, uc, Pb)
    r_o = rhs.copy(order="F")
    rp.compute_pseudo_pressure_rhs(U, V, W, r_o)
    rp.set_outflow_backflow_pressure(U, V, W, r_o, DT * c)
    plan = dfp.Plan(g)
    phib = np.zeros((g.nyg - 2, g.nzg - 2), order="F")
    plan.set_outflow(beta, uc, Pb, phib)
    r_d = rhs.copy(order="F")
    plan.divergence(U, V, W, r_d)
    plan.outflow_backflow(U, V, W, r_d, DT, c)
    assert rel_linf(phib, rp.phib) <= 1e-14
    assert rel_linf(r_d, r_o) <= 1e-14
    # the face correction with the same phi and phib (projection.f90:580-583)
    phi = np.asfortranarray(np.random.default_rng(3).standard_normal(g.shape_rhs))
    Uo, Vo, Wo = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    Ud, Vd, Wd = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    rp.phib = phib.copy(order="F")
    rp.project_velocity(Uo, Vo, Wo, phi)
    plan.correct(Ud, Vd, Wd, phi)
    assert np.array_equal(Uo, Ud) and np.array_equal(Vo, Vd) and np.array_equal(Wo, Wd)
    plan.close()


def test_inflow_device_resident_equals_host():
    g = io_grid(32, 24, 16, 2.0)
    U, V, W, P, rhs = io_fields(g, 99, 0.15)
    Pb, uc = backflow_state(g, 4)
    plan = dfp.Plan(g)
    h = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    ph_h = np.zeros((g.nyg - 2, g.nzg - 2), order="F")
    plan.set_outflow(0.5, uc, Pb, ph_h)
    plan.project(h[0], h[1], h[2], h[4], h[3], DT, C3, save_p=True)
    dev = [dfp.DeviceArray.from_host(a) for a in (U, V, W, P, rhs)]
    pb_d = dfp.DeviceArray.from_host(Pb)
    ph_d = dfp.DeviceArray((g.nyg - 2, g.nzg - 2), zero=True)
    plan.set_outflow(0.5, uc, pb_d, ph_d)
    plan.project(dev[0], dev[1], dev[2], dev[4], dev[3], DT, C3, save_p=True)
    for a, b in zip(h, dev):
        assert np.array_equal(a, b.download())
    assert np.array_equal(ph_h, ph_d.download())


@pytest.mark.parametrize("tuning", [(1, 0, 0, 1), (2, 0, 0, 1), (3, 0, 0, 1), (8, 0, 0, 1), (16, 0, 0, 1),
                                    (4, 0, 0, 2), (2, 2, 4, 1)],
                         ids=lambda t: f"R{t[0]}_T{t[1]}_G{t[2]}_f{t[3]}")
def test_inflow_z_splits_and_fused(tuning):
    """Every z split radix (the DCT-IV kernels are instantiated per radix),
    the fused z kernel and an explicit column/row shape."""
    g = io_grid(32, 12, 48, 1.6)
    print(f"inflow tuning {tuning}", flush=True)
    U, V, W, P, rhs = io_fields(g, 77, 0.1)
    o = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    op = ref_c.CProjection(g)
    op.compute_projection_step(o[0], o[1], o[2], o[4], o[3], DT, C3)
    op.close()
    plan = dfp.Plan(g, tuning=tuning)
    d = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    plan.project(d[0], d[1], d[2], d[4], d[3], DT, C3, save_p=True)
    plan.close()
    for lab, a, b in zip(("U", "V", "W", "P", "phi"), d, o):
        assert rel_linf(a, b) <= TOL, (tuning, lab)


@pytest.mark.parametrize("beta", [0.0, 0.6])
def test_inflow_full_size_parity(beta):
    """256x192x256 (nx = 257 faces), stretching 2.0, sigma 0.15, with and
    without the backflow term, and the divergence invariant."""
    g = io_grid(256, 192, 256, 2.0)
    print(f"inflow 256x192x256 beta {beta}", flush=True)
    o, d, fo, fd = run_both(g, 17, 0.15, beta)
    for lab, a, b in zip(("U", "V", "W", "P", "phi"), d, o):
        err = rel_linf(a, b)
        print(f"  {lab} rel-Linf {err:.3e}", flush=True)
        assert err <= TOL, lab
    if beta > 0:
        assert rel_linf(fd, fo) <= TOL
    U, V, W = d[0], d[1], d[2]
    dfp.refresh_velocity_ghosts(g, U, V, W)
    plan = dfp.Plan(g)
    md = plan.max_divergence(U, V, W)
    print(f"  max div {md:.3e} bound {div_bound(g, U, V, W):.3e}", flush=True)
    assert md <= div_bound(g, U, V, W)


def test_inflow_divergence_free_small():
    for beta in (0.0, 1.0):
        g = io_grid(24, 9, 18, 1.2)
        o, d, _, _ = run_both(g, 5, 0.2, beta)
        U, V, W = d[0], d[1], d[2]
        dfp.refresh_velocity_ghosts(g, U, V, W)
        plan = dfp.Plan(g)
        assert plan.max_divergence(U, V, W) <= div_bound(g, U, V, W), beta


def test_batch_refuses_inflow_plans():
    g = io_grid(16, 8, 8, 1.5)
    plan = dfp.Plan(g)
    from dfp.batch import Batch, Case

    U, V, W, P, rhs = (dfp.DeviceArray(s, zero=True) for s in (g.shape_u, g.shape_v, g.shape_w, g.shape_p,
                                                                g.shape_rhs))
    b = Batch(device=0, max_cases=4)
    with pytest.raises(dfp.DfpError) as ei:
        b.project([Case(plan, U, V, W, rhs, P, DT, C3, True)])
    assert ei.value.code == -2
    b.close()
    plan.close()
