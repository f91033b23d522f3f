"""GPU: the FP32+refine form of the projection against the FP64 path and the
C oracle on identical inputs.

Bars (written here, as the tests apply them):
  * after 1 refinement sweep (the library default), phi, P and the corrected
    U, V, W agree with the C oracle and with the FP64 GPU path to relative
    L-inf <= 1e-11;
  * the error falls with the sweep count (0 sweeps: the bare FP32 solve,
    relative error below 1e-4);
  * the form is deterministic: two runs on the same inputs are bit-identical.
"""
from __future__ import annotations

import numpy as np
import pytest

import dfp
from _cases import SMALL_CASES, div_bound, fields, make_grid, oracle_project, rel_linf

pytestmark = pytest.mark.gpu

MIX_TOL = 1e-11
DT = 1e-3
C3 = 3.0 / 4.0
LABELS = ("U", "V", "W", "P", "phi")


def _project(g, inputs, precision, sweeps=None, save_p=True):
    d = [a.copy(order="F") for a in inputs]
    plan = dfp.Plan(g, precision=precision, sweeps=sweeps)
    plan.project(d[0], d[1], d[2], d[4], d[3] if save_p else None, DT, C3, save_p=save_p)
    plan.close()
    return d


def _oracle(g, inputs):
    o = [a.copy(order="F") for a in inputs]
    oracle_project(g, *o, DT, C3, save_p=True)
    return o


@pytest.mark.parametrize("name,kw,sigma", SMALL_CASES, ids=[c[0] for c in SMALL_CASES])
def test_mixed_small_cases(name, kw, sigma):
    g = make_grid(**kw)
    print(f"case {name} FP32+refine", flush=True)
    inputs = fields(g, 2000 + len(name), sigma)
    o = _oracle(g, inputs)
    d = _project(g, inputs, dfp.DFP_FP32_REFINE)
    for lab, a, b in zip(LABELS, d, o):
        err = rel_linf(a, b)
        assert err <= MIX_TOL, f"{name} {lab}: rel-Linf {err:.3e}"


def test_mixed_error_falls_with_sweeps():
    g = make_grid(64, 64, 64, 1.5)
    print("case 64^3 FP32+refine sweeps 0..3", flush=True)
    inputs = fields(g, 2024, 0.1)
    o = _oracle(g, inputs)
    ref64 = _project(g, inputs, dfp.DFP_FP64)
    errs = []
    for s in range(4):
        d = _project(g, inputs, dfp.DFP_FP32_REFINE, sweeps=s)
        e_or = max(rel_linf(a, b) for a, b in zip(d, o))
        e_64 = max(rel_linf(a, b) for a, b in zip(d, ref64))
        print(f"  sweeps {s}: vs oracle {e_or:.3e}  vs FP64 GPU {e_64:.3e}", flush=True)
        errs.append(e_or)
    assert errs[0] < 1e-4
    assert errs[1] < errs[0]
    assert errs[1] <= MIX_TOL and errs[2] <= MIX_TOL and errs[3] <= MIX_TOL


def test_mixed_solve_from_rhs_and_skip_p_save():
    g = make_grid(48, 32, 24, 2.0)
    U, V, W, P, rhs = fields(g, 77, 0.1)
    plan64 = dfp.Plan(g)
    plan64.divergence(U, V, W, rhs)
    r64, P64 = rhs.copy(order="F"), P.copy(order="F")
    plan64.solve(r64, P64, DT, C3, save_p=True)
    r32, P32 = rhs.copy(order="F"), P.copy(order="F")
    plan32 = dfp.Plan(g, precision=dfp.DFP_FP32_REFINE)
    plan32.solve(r32, P32, DT, C3, save_p=True)
    assert rel_linf(r32, r64) <= MIX_TOL
    assert rel_linf(P32, P64) <= MIX_TOL
    Pk = np.full(g.shape_p, 7.0, order="F")
    r32b = rhs.copy(order="F")
    plan32.solve(r32b, Pk, DT, C3, save_p=False)
    assert np.all(Pk == 7.0)
    assert np.array_equal(r32b, r32)


def test_mixed_is_deterministic():
    g = make_grid(64, 48, 32, 2.0)
    inputs = fields(g, 5, 0.15)
    a = _project(g, inputs, dfp.DFP_FP32_REFINE)
    b = _project(g, inputs, dfp.DFP_FP32_REFINE)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_mixed_sweep_bounds():
    g = make_grid(16, 8, 8, 1.5)
    plan = dfp.Plan(g, precision=dfp.DFP_FP32_REFINE)
    for bad in (-1, 9):
        with pytest.raises(dfp.DfpError):
            plan.set_refine_sweeps(bad)
    plan.set_refine_sweeps(8)


def test_mixed_config2_size():
    # 256x192x256, stretching 2.0, sigma 0.15
    g = make_grid(256, 192, 256, 2.0)
    print("case 256x192x256 FP32+refine", flush=True)
    inputs = fields(g, 1, 0.15)
    o = _oracle(g, inputs)
    ref64 = _project(g, inputs, dfp.DFP_FP64)
    for s in (1, 2, 3):
        d = _project(g, inputs, dfp.DFP_FP32_REFINE, sweeps=s)
        e_or = {lab: rel_linf(a, b) for lab, a, b in zip(LABELS, d, o)}
        e_64 = {lab: rel_linf(a, b) for lab, a, b in zip(LABELS, d, ref64)}
        print(f"  sweeps {s}: phi vs oracle {e_or['phi']:.3e} vs FP64 {e_64['phi']:.3e}; "
              f"max over fields {max(e_or.values()):.3e} / {max(e_64.values()):.3e}", flush=True)
        assert max(e_or.values()) <= MIX_TOL and max(e_64.values()) <= MIX_TOL
    U, V, W = d[0], d[1], d[2]
    dfp.refresh_velocity_ghosts(g, U, V, W)
    md = dfp.Plan(g).max_divergence(U, V, W)
    print(f"  max div after FP32+refine {md:.3e} (FP64 bar {div_bound(g, U, V, W):.3e})", flush=True)


def test_mixed_config3_size():
    # 512x384x512, stretching 2.0, sigma 0.15: sweeps 1..3 against the C oracle and the FP64 GPU path
    g = make_grid(512, 384, 512, 2.0)
    print("case 512x384x512 FP32+refine", flush=True)
    inputs = fields(g, 3, 0.15)
    o = _oracle(g, inputs)
    ref64 = _project(g, inputs, dfp.DFP_FP64)
    for s in (1, 2, 3):
        d = _project(g, inputs, dfp.DFP_FP32_REFINE, sweeps=s)
        e_or = max(rel_linf(a, b) for a, b in zip(d, o))
        e_64 = max(rel_linf(a, b) for a, b in zip(d, ref64))
        print(f"  sweeps {s}: max over fields vs oracle {e_or:.3e}, vs FP64 GPU {e_64:.3e}", flush=True)
        assert e_or <= MIX_TOL and e_64 <= MIX_TOL
        del d
