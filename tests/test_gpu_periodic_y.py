"""GPU parity of the periodic-y branch (y_bc_type = 0 on a uniform y grid with
periodic x and z: the y FFT over the nym-1 independent rows in place of the
wall solve, projection.f90:246-273, the duplicate row :350-354, and the
periodic y wrap of apply_boundary_conditions, boundary_conditions.f90:229-233,
:762-781) against the C oracle on identical inputs.

Bars (written here, as the tests apply them):
  * phi (rhs_p, duplicates included), P and
    the corrected U, V, W: relative L-inf <= 1e-12
  * post-projection max|div u| <= 1e-12 * max|u| / min spacing, after the
    periodic ghosts of the velocity are refreshed in x, y and z
  * the boundary conditions: bit-identical (same copies)
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

# (name, (nxp, nym, nzp)): nym - 1 independent y rows, 2^a 3^b
PY_CASES = [
    ("tiny_8x5x6", (8, 5, 6)),
    ("radix3_24x10x18", (24, 10, 18)),
    ("asym_12x13x8", (12, 13, 8)),
    ("nym2_16x2x16", (16, 2, 16)),
    ("nzp1_6x7x1", (6, 7, 1)),
    ("cube_64x65x64", (64, 65, 64)),
    ("tall_32x129x16", (32, 129, 16)),
]


def py_fields(g, seed, sigma=0.3):
    """A mean flow in x plus Gaussian noise on every component, every ghost
    and duplicate plane set periodic in x, y and z (the state
    apply_boundary_conditions leaves), and an rhs_p that is random except at
    its ghost positions i = nxg and k = nzg: the reference zeroes rhs_p at
    allocation (initialization.f90:220) and never writes those positions from
    anything but each other, so they hold zero throughout its runs."""
    rng = np.random.default_rng(seed)
    U = np.asfortranarray(1.0 + sigma * rng.standard_normal(g.shape_u))
    V = np.asfortranarray(sigma * rng.standard_normal(g.shape_v))
    W = np.asfortranarray(sigma * rng.standard_normal(g.shape_w))
    R.apply_boundary_conditions(g, U, V, W, None, None, None)
    P = np.zeros(g.shape_p, order="F")
    rhs = np.asfortranarray(rng.standard_normal(g.shape_rhs))
    rhs[-1, :, :] = 0.0
    rhs[:, :, -1] = 0.0
    return U, V, W, P, rhs


def run_both(g, seed, c=C3, save_p=True, tuning=None):
    U, V, W, P, rhs = py_fields(g, seed)
    o = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    op = ref_c.CProjection(g)
    op.compute_projection_step(o[0], o[1], o[2], o[4], o[3] if save_p else None, DT, c)
    op.close()
    plan = dfp.Plan(g, tuning=tuning)
    d = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    plan.project(d[0], d[1], d[2], d[4], d[3] if save_p else None, DT, c, save_p=save_p)
    plan.close()
    return o, d


def _divergence_after_refresh(g, U, V, W):
    U, V, W = (a.copy(order="F") for a in (U, V, W))
    R.apply_boundary_conditions(g, U, V, W, None, None, None)
    op = ref_c.CProjection(g)
    md = op.check_divergence(U, V, W)
    op.close()
    return md, div_bound(g, U, V, W)


@pytest.mark.parametrize("name,dims", PY_CASES, ids=[c[0] for c in PY_CASES])
def test_periodic_y_project_small_cases(name, dims):
    g = dfp.periodic_y_grid(*dims)
    print(f"periodic-y case {name}", flush=True)
    o, d = run_both(g, 500 + dims[0] + dims[1])
    for lab, a, b in zip(("U", "V", "W", "P", "phi"), d, o):
        err = rel_linf(a, b)
        assert err <= TOL, f"{name} {lab}: rel-Linf {err:.3e}"
    # the duplicate row equals row 0 over the whole x-z extent, bit for bit
    assert np.array_equal(d[4][:, g.nym - 1, :], d[4][:, 0, :])
    md, bound = _divergence_after_refresh(g, *d[:3])
    assert md <= bound, (name, md, bound)


def test_periodic_y_solve_and_device_resident():
    g = dfp.periodic_y_grid(32, 25, 24)
    U, V, W, P, rhs = py_fields(g, 9)
    op = ref_c.CProjection(g)
    plan = dfp.Plan(g)
    r_o = rhs.copy(order="F"); r_d = rhs.copy(order="F")
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
    plan.close()


@pytest.mark.parametrize("tuning", [(1, 0, 0, 2), (2, 0, 0, 2), (3, 0, 0, 2), (4, 0, 0, 2), (8, 0, 0, 2),
                                    (16, 0, 0, 2), (4, 2, 0, 0)],
                         ids=lambda t: f"R{t[0]}_T{t[1]}_f{t[3]}")
def test_periodic_y_z_splits(tuning):
    g = dfp.periodic_y_grid(32, 17, 48)
    print(f"periodic-y tuning {tuning}", flush=True)
    o, d = run_both(g, 41, tuning=tuning)
    for lab, a, b in zip(("U", "V", "W", "P", "phi"), d, o):
        assert rel_linf(a, b) <= TOL, (tuning, lab)


def test_periodic_y_refuses_the_split_z_kernels():
    g = dfp.periodic_y_grid(16, 9, 16)
    with pytest.raises(dfp.DfpError) as ei:
        dfp.Plan(g, tuning=(4, 0, 0, 1))
    assert ei.value.code == -2


def test_periodic_y_boundary_conditions_and_substep():
    g = dfp.periodic_y_grid(16, 13, 12)
    rng = np.random.default_rng(3)
    F3 = [np.asfortranarray(rng.standard_normal(s)) for s in (g.shape_u, g.shape_v, g.shape_w)]
    o = [a.copy(order="F") for a in F3]
    op = ref_c.CProjection(g)
    op.apply_boundary_conditions(*o, None, None, None)
    plan = dfp.Plan(g)
    d = [dfp.DeviceArray.from_host(a) for a in F3]
    plan.apply_bc(*d)
    for x, y in zip(d, o):
        assert np.array_equal(x.download(), y)
    # one RK3 substage (time_integration.f90:256-304) on the periodic box
    stage = 2
    coef = np.array([1 / 4, 5 / 12, 0.0])
    Uo = [np.asfortranarray(rng.standard_normal(s)) for s in (g.shape_u, g.shape_v, g.shape_w)]
    inner = ((g.nx - 2, g.nyg - 2, g.nzg - 2), (g.nxg - 2, g.ny - 2, g.nzg - 2), (g.nxg - 2, g.nyg - 2, g.nz - 2))
    Fr = [[np.asfortranarray(rng.standard_normal(sh)) for _ in range(3)] for sh in inner]
    P = np.zeros(g.shape_p, order="F")
    rhs = np.zeros(g.shape_rhs, order="F")
    o = [a.copy(order="F") for a in (*F3, P, rhs)]
    op.rk_update(stage, o[0], o[1], o[2], *Uo, *Fr, coef, DT)
    op.apply_boundary_conditions(o[0], o[1], o[2], None, None, None)
    op.compute_projection_step(o[0], o[1], o[2], o[4], None, DT, coef[stage - 1])
    op.apply_boundary_conditions(o[0], o[1], o[2], None, None, None)
    op.close()
    d = [dfp.DeviceArray.from_host(a) for a in (*F3, P, rhs)]
    plan.substep(stage, d[0], d[1], d[2], *[dfp.DeviceArray.from_host(a) for a in Uo],
                 *[[dfp.DeviceArray.from_host(x) for x in f] for f in Fr], coef, None, None, None, d[4], None, DT)
    for lab, x, y in zip(("U", "V", "W", "P", "phi"), d, o):
        assert rel_linf(x.download(), y) <= TOL, lab
    plan.close()


def test_periodic_y_full_size_parity():
    """A 256^3 periodic box (nym = 257: the last row duplicates the first)."""
    g = dfp.periodic_y_grid(256, 257, 256)
    print("periodic-y 256x257x256", flush=True)
    o, d = run_both(g, 17)
    for lab, a, b in zip(("U", "V", "W", "P", "phi"), d, o):
        err = rel_linf(a, b)
        assert err <= TOL, f"{lab}: rel-Linf {err:.3e}"
    md, bound = _divergence_after_refresh(g, *d[:3])
    assert md <= bound, (md, bound)
