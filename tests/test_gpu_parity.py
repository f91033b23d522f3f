"""GPU parity: the HIP projection against the C oracle on identical inputs.

Bars (written here, as the tests apply them):
  * divergence, correction, max-divergence: bit-identical (same operations)
  * phi (rhs_p), P and the corrected U, V, W: relative L-inf <= 1e-12
  * post-projection max|div u| <= 1e-12 * max|u| / min spacing (after the
    periodic velocity ghosts are refreshed, as the reference's monitor sees them)
"""
from __future__ import annotations

import glob
import os

import numpy as np
import pytest

import dfp
from oracle import ref_c
from _cases import (SMALL_CASES, chan_les_small_grid, div_bound, fields, make_grid, oracle_project, rel_linf)

pytestmark = pytest.mark.gpu

TOL = 1e-12
DT = 1e-3
C3 = 3.0 / 4.0


def _run_both(g, seed, sigma, save_p=True):
    U, V, W, P, rhs = fields(g, seed, sigma)
    o = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    oracle_project(g, *o, DT, C3, save_p=save_p)
    plan = dfp.Plan(g)
    d = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    plan.project(d[0], d[1], d[2], d[4], d[3] if save_p else None, DT, C3, save_p=save_p)
    plan.close()
    return o, d


@pytest.mark.parametrize("name,kw,sigma", SMALL_CASES, ids=[c[0] for c in SMALL_CASES])
def test_project_small_cases(name, kw, sigma):
    g = make_grid(**kw)
    print(f"case {name}", flush=True)
    o, d = _run_both(g, 1000 + len(name), sigma)
    for lab, a, b in zip(("U", "V", "W", "P", "phi"), d, o):
        err = rel_linf(a, b)
        assert err <= TOL, f"{name} {lab}: rel-Linf {err:.3e}"


def test_project_chan_les_small_geometry():
    g = chan_les_small_grid()
    o, d = _run_both(g, 7, 0.1)
    for lab, a, b in zip(("U", "V", "W", "P", "phi"), d, o):
        assert rel_linf(a, b) <= TOL, lab


def test_stages_bit_identical_and_solve():
    g = make_grid(24, 16, 12, 1.5)
    U, V, W, P, rhs = fields(g, 11, 0.1)
    op = ref_c.CProjection(g)
    plan = dfp.Plan(g)
    # divergence (projection.f90:37-63)
    r_o = rhs.copy(order="F"); r_d = rhs.copy(order="F")
    op.compute_pseudo_pressure_rhs(U, V, W, r_o)
    plan.divergence(U, V, W, r_d)
    assert np.array_equal(r_o, r_d)
    # solve (projection.f90:100-382) with P
    P_o = P.copy(order="F"); P_d = P.copy(order="F")
    op.solve_poisson_equation(r_o, P_o, DT, C3)
    plan.solve(r_d, P_d, DT, C3, save_p=True)
    assert rel_linf(r_d, r_o) <= TOL
    assert rel_linf(P_d, P_o) <= TOL
    # correction from the same phi (projection.f90:551-604)
    Uo, Vo, Wo = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    Ud, Vd, Wd = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    op.project_velocity(Uo, Vo, Wo, r_o)
    plan.correct(Ud, Vd, Wd, r_o)
    assert np.array_equal(Uo, Ud) and np.array_equal(Vo, Vd) and np.array_equal(Wo, Wd)
    # max divergence (projection.f90:607-639)
    assert plan.max_divergence(U, V, W) == op.check_divergence(U, V, W)
    op.close(); plan.close()


def test_skip_p_save_leaves_P():
    g = make_grid(16, 8, 8, 1.5)
    U, V, W, P, rhs = fields(g, 5, 0.1)
    P[:] = 123.0
    plan = dfp.Plan(g)
    plan.project(U, V, W, rhs, P, DT, C3, save_p=False)
    assert np.all(P == 123.0)


def test_divergence_invariant_config1():
    # 64^3, stretching 1.5, sigma 0.1 (the plain single-projection case)
    g = make_grid(64, 64, 64, 1.5)
    print("case config1 64^3", flush=True)
    o, d = _run_both(g, 2024, 0.1)
    for lab, a, b in zip(("U", "V", "W", "P", "phi"), d, o):
        assert rel_linf(a, b) <= TOL, lab
    U, V, W = d[0], d[1], d[2]
    dfp.refresh_velocity_ghosts(g, U, V, W)
    plan = dfp.Plan(g)
    assert plan.max_divergence(U, V, W) <= div_bound(g, U, V, W)


def test_device_resident_equals_host():
    g = make_grid(32, 24, 16, 2.0)
    U, V, W, P, rhs = fields(g, 99, 0.15)
    plan = dfp.Plan(g)
    h = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    plan.project(h[0], h[1], h[2], h[4], h[3], DT, C3, save_p=True)
    dev = [dfp.DeviceArray.from_host(a) for a in (U, V, W, P, rhs)]
    plan.project(dev[0], dev[1], dev[2], dev[4], dev[3], DT, C3, save_p=True)
    for a, b in zip(h, dev):
        assert np.array_equal(a, b.download())


def test_device_synth_matches_host_generator():
    g = make_grid(16, 12, 8, 1.5)
    plan = dfp.Plan(g)
    dev = [dfp.DeviceArray(s) for s in (g.shape_u, g.shape_v, g.shape_w)]
    plan.synth(dev[0], dev[1], dev[2], 31337, 0.15)
    host = dfp.synth_fields(g, 31337, 0.15)
    for a, b in zip(host, dev):
        np.testing.assert_allclose(b.download(), a, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("seed", [1, 2])
def test_config2_full_size_parity(seed):
    # 256x192x256, stretching 2.0, sigma 0.15: full-size parity with the C oracle
    g = make_grid(256, 192, 256, 2.0)
    print(f"case config2 256x192x256 seed {seed}", flush=True)
    o, d = _run_both(g, seed, 0.15)
    for lab, a, b in zip(("U", "V", "W", "P", "phi"), d, o):
        err = rel_linf(a, b)
        print(f"  {lab} rel-Linf {err:.3e}", flush=True)
        assert err <= TOL, lab
    U, V, W = d[0], d[1], d[2]
    dfp.refresh_velocity_ghosts(g, U, V, W)
    plan = dfp.Plan(g)
    md = plan.max_divergence(U, V, W)
    print(f"  max div {md:.3e} bound {div_bound(g, U, V, W):.3e}", flush=True)
    assert md <= div_bound(g, U, V, W)


# projection fixtures (the snapshot-reader fixture has its own test)
GOLDEN = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
                if not os.path.basename(p).startswith("snapshot_"))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_gpu_matches_golden_fixtures(path):
    """The HIP path against the numpy oracle's stored outputs (independent of the C oracle)."""
    f = np.load(path)
    nx, ny, nz = (int(v) for v in f["namelist"])
    Lx, Ly, Lz = (float(v) for v in f["lengths"])
    g = dfp.namelist_grid(nx, ny, nz, Lx, Ly, Lz, float(f["alpha"]), int(f["grid_type"]))
    U, V, W = (np.asfortranarray(f[k]) for k in ("U_in", "V_in", "W_in"))
    P = np.zeros(g.shape_p, order="F")
    rhs = np.zeros(g.shape_rhs, order="F")
    plan = dfp.Plan(g)
    plan.project(U, V, W, rhs, P, float(f["dt"]), float(f["c"]), save_p=True)
    for lab, x, key in (("U", U, "U_out"), ("V", V, "V_out"), ("W", W, "W_out"), ("P", P, "P_out"), ("phi", rhs, "rhs_out")):
        assert rel_linf(x, f[key]) <= TOL, lab


@pytest.mark.parametrize("nxp,nym,nzp", [(512, 10, 192), (576, 6, 144)])
def test_wide_x_strip_order(nxp, nym, nzp):
    """From nxp = 512 the x kernels walk their tiles in strips of 8 column
    blocks (12 and 36 blocks here, so the last strip is partial): every tile
    is still visited once, FP64 and FP32+refine against the C oracle."""
    g = make_grid(nxp, nym, nzp, 1.8)
    print(f"wide x {nxp}x{nym}x{nzp}", flush=True)
    U, V, W, P, rhs = fields(g, nxp + nzp, 0.1)
    o = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    oracle_project(g, *o, DT, C3, save_p=True)
    for prec, tol in ((dfp.DFP_FP64, TOL), (dfp.DFP_FP32_REFINE, 1e-11)):
        plan = dfp.Plan(g, precision=prec)
        d = [a.copy(order="F") for a in (U, V, W, P, rhs)]
        plan.project(d[0], d[1], d[2], d[4], d[3], DT, C3, save_p=True)
        plan.close()
        for lab, a, b in zip(("U", "V", "W", "P", "phi"), d, o):
            assert rel_linf(a, b) <= tol, (prec, lab)


def _sweep_shapes(n, seed):
    """n seeded grid shapes with 2^a 3^b extents: nxp 4..576 (even), nym 2..80,
    nzp 1..288, stretching 0.5..2.5; wide ones are kept thin so the oracle
    stays quick."""
    rng = np.random.default_rng(seed)
    ext = sorted({2 ** a * 3 ** b for a in range(0, 10) for b in range(0, 3) if 2 ** a * 3 ** b <= 576})
    out = []
    while len(out) < n:
        nxp = int(rng.choice([e for e in ext if e % 2 == 0 and e >= 4]))
        nzp = int(rng.choice([e for e in ext if e <= 288]))
        nym = int(rng.integers(2, 81))
        if nxp * nym * nzp > 600_000:
            continue
        out.append((nxp, nym, nzp, float(rng.uniform(0.5, 2.5))))
    return out


@pytest.mark.parametrize("shape", _sweep_shapes(24, 2026), ids=lambda s: f"{s[0]}x{s[1]}x{s[2]}")
def test_seeded_shape_sweep(shape):
    """Seeded sweep over grid shapes (z splits, row lengths, column counts and
    x tile rules the fixed cases do not reach): FP64 to 1e-12 and
    FP32+refine to 1e-11 against the C oracle."""
    nxp, nym, nzp, alpha = shape
    g = make_grid(nxp, nym, nzp, alpha)
    print(f"sweep {nxp}x{nym}x{nzp} alpha {alpha:.3f}", flush=True)
    U, V, W, P, rhs = fields(g, nxp * 7 + nym * 3 + nzp, 0.1)
    o = [a.copy(order="F") for a in (U, V, W, P, rhs)]
    oracle_project(g, *o, DT, C3, save_p=True)
    for prec, tol in ((dfp.DFP_FP64, TOL), (dfp.DFP_FP32_REFINE, 1e-11)):
        plan = dfp.Plan(g, precision=prec)
        d = [a.copy(order="F") for a in (U, V, W, P, rhs)]
        plan.project(d[0], d[1], d[2], d[4], d[3], DT, C3, save_p=True)
        plan.close()
        for lab, a, b in zip(("U", "V", "W", "P", "phi"), d, o):
            err = rel_linf(a, b)
            assert err <= tol, f"{shape} prec {prec} {lab}: {err:.3e}"


@pytest.mark.parametrize("field,value", [("U", np.nan), ("V", np.inf), ("W", -2.0e8)])
def test_check_finite_refuses_bad_fields(field, value):
    """The monitor's guard (monitor.f90:32-38) as an opt-in of dfp_project: a
    NaN, an Inf or |u| > 1e8 in u* is refused with DFP_E_NONFINITE before
    anything is written, host-staged and device-resident; a clean field
    projects as without the guard."""
    g = make_grid(16, 10, 8, 1.5)
    U, V, W, P, rhs = fields(g, 8, 0.1)
    bad = [a.copy(order="F") for a in (U, V, W)]
    Fb = bad["UVW".index(field)]
    Fb[3, 4, 5] = value
    plan = dfp.Plan(g)
    rhs0 = rhs.copy(order="F")
    with pytest.raises(dfp.DfpError) as ei:
        plan.project(bad[0], bad[1], bad[2], rhs0, P, DT, C3, save_p=True, check_finite=True)
    assert ei.value.code == -5
    assert np.array_equal(rhs0, rhs) and np.isnan(Fb[3, 4, 5]) == np.isnan(value)
    dev = [dfp.DeviceArray.from_host(a) for a in (*bad, P, rhs)]
    with pytest.raises(dfp.DfpError):
        plan.project(dev[0], dev[1], dev[2], dev[4], dev[3], DT, C3, save_p=True, check_finite=True)
    assert np.array_equal(dev[4].download(), rhs)
    # a value outside the monitor's ranges (a ghost) is not refused; a clean field projects
    ok = [a.copy(order="F") for a in (U, V, W)]
    ok[0][0, 0, 0] = np.nan
    a = [x.copy(order="F") for x in (*ok, P, rhs)]
    b = [x.copy(order="F") for x in (*ok, P, rhs)]
    plan.project(a[0], a[1], a[2], a[4], a[3], DT, C3, save_p=True, check_finite=True)
    plan.project(b[0], b[1], b[2], b[4], b[3], DT, C3, save_p=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y, equal_nan=True)
    plan.close()
