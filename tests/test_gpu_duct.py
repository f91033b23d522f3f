"""GPU parity of the 4-wall duct branch (x periodic, y walls, z walls: the z
eigenbasis transforms and per-mode y solves of projection.f90:172-218, set-up
initialization.f90:544-618) against the C oracle on identical inputs.

Bars (written here, as the tests apply them):
  * phi (rhs_p), P and the corrected U, V, W: relative L-inf <= 1e-12 (the
    GPU's eigenbasis comes from its own implicit-QL eigensolver, the C
    oracle's from a Jacobi iteration: the two agree to round-off)
  * post-projection max|div u| <= 1e-12 * max|u| / min spacing, for inputs
    with no flux through the walls, after the periodic x ghosts are
    refreshed
"""
from __future__ import annotations

import numpy as np
import pytest

import dfp
from oracle import ref_c
from _cases import div_bound, rel_linf

pytestmark = pytest.mark.gpu

TOL = 1e-12
DT = 1e-3
C3 = 3.0 / 4.0

# (name, (nxp, nym, nzm, alpha_y, alpha_z))
DUCT_CASES = [
    ("tiny_8x6x5", (8, 6, 5, 1.5, 0.0)),
    ("stretched_6x7x8", (6, 7, 8, 1.2, 1.8)),
    ("uniform_4x5x3", (4, 5, 3, 0.0, 0.0)),
    ("asym_12x4x6", (12, 4, 6, 2.0, 2.5)),
    ("one_z_cell_16x12x1", (16, 12, 1, 1.5, 0.0)),
    ("odd_32x33x65", (32, 33, 65, 1.5, 1.5)),
    ("square_64x48x64", (64, 48, 64, 2.0, 1.0)),
    ("wide_96x20x130", (96, 20, 130, 1.0, 2.0)),
]


def _refresh_x(g, U, V, W):
    U[0] = U[g.nx - 2]; U[g.nx - 1] = U[1]
    for F in (V, W):
        F[0] = F[g.nxg - 3]; F[g.nxg - 2] = F[1]; F[g.nxg - 1] = F[2]


def duct_fields(g, seed, sigma=0.3):
    """A mean flow in x plus Gaussian noise, no flux through the walls (V at
    the y walls and W at the z walls zero), periodic x ghosts, rhs_p zero
    (as the reference allocates it, initialization.f90:220)."""
    rng = np.random.default_rng(seed)
    U = np.asfortranarray(1.0 + sigma * rng.standard_normal(g.shape_u))
    V = np.asfortranarray(sigma * rng.standard_normal(g.shape_v))
    W = np.asfortranarray(sigma * rng.standard_normal(g.shape_w))
    V[:, 0, :] = 0.0; V[:, -1, :] = 0.0
    W[:, :, 0] = 0.0; W[:, :, -1] = 0.0
    _refresh_x(g, U, V, W)
    return U, V, W, np.zeros(g.shape_p, order="F"), np.zeros(g.shape_rhs, order="F")


def run_both(g, seed, save_p=True):
    U, V, W, P, rhs = duct_fields(g, seed)
    o = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    op = ref_c.CProjection(g)
    op.compute_projection_step(o[0], o[1], o[2], o[4], o[3] if save_p else None, DT, C3)
    op.close()
    plan = dfp.Plan(g)
    d = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    plan.project(d[0], d[1], d[2], d[4], d[3] if save_p else None, DT, C3, save_p=save_p)
    plan.close()
    return o, d


def _check(g, o, d, name):
    for lab, a, b in zip(("U", "V", "W", "P", "phi"), d, o):
        err = rel_linf(a, b)
        assert err <= TOL, f"{name} {lab}: rel-Linf {err:.3e}"
    if g.nzp == 1:
        # one cell between the z walls: the reference's Dzz(2,2) is assigned
        # twice (b+c, then a+b, initialization.f90:559-572), which leaves -c
        # where the Neumann operator has 0, so its projection (and this one,
        # equal to it above) is not divergence-free there
        return
    U, V, W = (a.copy(order="F") for a in d[:3])
    _refresh_x(g, U, V, W)
    op = ref_c.CProjection(g)
    md = op.check_divergence(U, V, W)
    op.close()
    assert md <= div_bound(g, U, V, W), (name, md)


@pytest.mark.parametrize("name,dims", DUCT_CASES, ids=[c[0] for c in DUCT_CASES])
def test_duct_project_small_cases(name, dims):
    g = dfp.duct_grid(*dims)
    print(f"duct case {name}", flush=True)
    o, d = run_both(g, 900 + dims[0] + dims[2])
    _check(g, o, d, name)


def test_duct_solve_and_device_resident():
    g = dfp.duct_grid(24, 16, 20, 1.7, 1.3)
    U, V, W, P, rhs = duct_fields(g, 12)
    rng = np.random.default_rng(5)
    r = rhs.copy(order="F")
    r[:g.nxp, :, :g.nzp] = rng.standard_normal((g.nxp, g.nym, g.nzp))
    op = ref_c.CProjection(g)
    plan = dfp.Plan(g)
    r_o = r.copy(order="F"); r_d = r.copy(order="F")
    P_o = P.copy(order="F"); P_d = P.copy(order="F")
    op.solve_poisson_equation(r_o, P_o, DT, C3)
    plan.solve(r_d, P_d, DT, C3, save_p=True)
    assert rel_linf(r_d, r_o) <= TOL and rel_linf(P_d, P_o) <= TOL
    op.close()
    h = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    plan.project(h[0], h[1], h[2], h[4], h[3], DT, C3, save_p=True)
    dev = [dfp.DeviceArray.from_host(a) for a in (U, V, W, P, rhs)]
    plan.project(dev[0], dev[1], dev[2], dev[4], dev[3], DT, C3, save_p=True)
    for a, b in zip(h, dev):
        assert np.array_equal(a, b.download())
    assert plan.tuning[0] == 1   # no z split
    plan.close()


def test_duct_refuses_a_z_split():
    g = dfp.duct_grid(16, 8, 12, 1.5)
    with pytest.raises(dfp.DfpError) as ei:
        dfp.Plan(g, tuning=(4, 0, 0, 0))  # the duct has no z FFT to split
    assert ei.value.code == -2


def test_duct_boundary_conditions_and_substep():
    """apply_boundary_conditions with the z walls (boundary_conditions.f90:
    215-223, :737-758) bit-identical, and one RK3 substage of the duct
    (time_integration.f90:256-304) to 1e-12."""
    g = dfp.duct_grid(16, 10, 12, 1.5, 1.2)
    rng = np.random.default_rng(8)
    F3 = [np.asfortranarray(rng.standard_normal(s)) for s in (g.shape_u, g.shape_v, g.shape_w)]
    al = (np.asfortranarray(rng.uniform(0, 0.2, (g.nx, 2, g.nzg))),
          np.asfortranarray(rng.uniform(0, 0.2, (g.nxg, 2, g.nzg))),
          np.asfortranarray(rng.uniform(0, 0.2, (g.nxg, 2, g.nz))))
    op = ref_c.CProjection(g)
    o = [a.copy(order="F") for a in F3]
    op.apply_boundary_conditions(*o, *al)
    plan = dfp.Plan(g)
    d = [dfp.DeviceArray.from_host(a) for a in F3]
    dal = [dfp.DeviceArray.from_host(a) for a in al]
    plan.apply_bc(*d, *dal)
    for x, y in zip(d, o):
        assert np.array_equal(x.download(), y)
    stage = 3
    coef = np.array([1 / 4, 0.0, 3 / 4])
    Uo = [np.asfortranarray(rng.standard_normal(s)) for s in (g.shape_u, g.shape_v, g.shape_w)]
    inner = ((g.nx - 2, g.nyg - 2, g.nzg - 2), (g.nxg - 2, g.ny - 2, g.nzg - 2), (g.nxg - 2, g.nyg - 2, g.nz - 2))
    Fr = [[np.asfortranarray(rng.standard_normal(sh)) for _ in range(3)] for sh in inner]
    P = np.zeros(g.shape_p, order="F")
    rhs = np.zeros(g.shape_rhs, order="F")
    o = [a.copy(order="F") for a in (*F3, P, rhs)]
    op.rk_update(stage, o[0], o[1], o[2], *Uo, *Fr, coef, DT)
    op.apply_boundary_conditions(o[0], o[1], o[2], *al)
    op.compute_projection_step(o[0], o[1], o[2], o[4], o[3], DT, coef[stage - 1])
    op.apply_boundary_conditions(o[0], o[1], o[2], *al)
    op.close()
    d = [dfp.DeviceArray.from_host(a) for a in (*F3, P, rhs)]
    plan.substep(stage, d[0], d[1], d[2], *[dfp.DeviceArray.from_host(a) for a in Uo],
                 *[[dfp.DeviceArray.from_host(x) for x in f] for f in Fr], coef, *dal, d[4], d[3], DT, save_p=True)
    for lab, x, y in zip(("U", "V", "W", "P", "phi"), d, o):
        assert rel_linf(x.download(), y) <= TOL, lab
    plan.close()


def test_duct_full_size_parity():
    """256x192x256 cells with y and z walls (stretching 2.0 both)."""
    g = dfp.duct_grid(256, 192, 256, 2.0, 2.0)
    print("duct 256x192x256", flush=True)
    o, d = run_both(g, 23)
    _check(g, o, d, "duct256")


# spanwise walls with a periodic y: (name, (nxp, nym, nzm, alpha_z))
ZWPY_CASES = [
    ("tiny_8x5x6", (8, 5, 6, 0.0)),
    ("stretched_6x9x7", (6, 9, 7, 1.5)),
    ("radix3_12x10x5", (12, 10, 5, 2.0)),
    ("nym2_8x2x4", (8, 2, 4, 1.0)),
    ("big_64x65x48", (64, 65, 48, 1.2)),
    ("wide_96x33x130", (96, 33, 130, 2.0)),
]


def zwpy_fields(g, seed, sigma=0.3):
    """As duct_fields with the periodic y: V, U and W periodic in y, W zero
    on the z walls."""
    from oracle import ref_numpy as R

    rng = np.random.default_rng(seed)
    U = np.asfortranarray(1.0 + sigma * rng.standard_normal(g.shape_u))
    V = np.asfortranarray(sigma * rng.standard_normal(g.shape_v))
    W = np.asfortranarray(sigma * rng.standard_normal(g.shape_w))
    W[:, :, 0] = 0.0; W[:, :, -1] = 0.0
    R.apply_boundary_conditions(g, U, V, W, None, None, None)
    return U, V, W, np.zeros(g.shape_p, order="F"), np.zeros(g.shape_rhs, order="F")


@pytest.mark.parametrize("name,dims", ZWPY_CASES, ids=[c[0] for c in ZWPY_CASES])
def test_zwall_periodic_y_project(name, dims):
    from oracle import ref_numpy as R

    g = dfp.zwall_periodic_y_grid(*dims)
    print(f"z walls, periodic y: {name}", flush=True)
    U, V, W, P, rhs = zwpy_fields(g, 70 + dims[0])
    o = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    op = ref_c.CProjection(g)
    op.compute_projection_step(o[0], o[1], o[2], o[4], o[3], DT, C3)
    plan = dfp.Plan(g)
    d = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    plan.project(d[0], d[1], d[2], d[4], d[3], DT, C3, save_p=True)
    plan.close()
    for lab, a, b in zip(("U", "V", "W", "P", "phi"), d, o):
        err = rel_linf(a, b)
        assert err <= TOL, f"{name} {lab}: rel-Linf {err:.3e}"
    # the duplicate row carries row 0
    assert np.max(np.abs(d[4][:, g.nym - 1, :] - d[4][:, 0, :])) <= 1e-14 * np.max(np.abs(d[4]))
    U2, V2, W2 = (a.copy(order="F") for a in d[:3])
    R.apply_boundary_conditions(g, U2, V2, W2, None, None, None)
    md = op.check_divergence(U2, V2, W2)
    op.close()
    assert md <= div_bound(g, U2, V2, W2), (name, md)


def test_zwall_periodic_y_full_size_parity():
    """256 x 193 x 256 (192 independent y rows), z faces stretched 2.0."""
    from oracle import ref_numpy as R

    g = dfp.zwall_periodic_y_grid(256, 193, 256, 2.0)
    print("z walls, periodic y: 256x193x256", flush=True)
    U, V, W, P, rhs = zwpy_fields(g, 31)
    o = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    op = ref_c.CProjection(g)
    op.compute_projection_step(o[0], o[1], o[2], o[4], o[3], DT, C3)
    plan = dfp.Plan(g)
    d = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    plan.project(d[0], d[1], d[2], d[4], d[3], DT, C3, save_p=True)
    plan.close()
    for lab, a, b in zip(("U", "V", "W", "P", "phi"), d, o):
        assert rel_linf(a, b) <= TOL, lab
    U2, V2, W2 = (a.copy(order="F") for a in d[:3])
    R.apply_boundary_conditions(g, U2, V2, W2, None, None, None)
    md = op.check_divergence(U2, V2, W2)
    op.close()
    assert md <= div_bound(g, U2, V2, W2), md
