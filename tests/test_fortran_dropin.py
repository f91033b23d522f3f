"""The Fortran drop-in module (dopamine-fdm_amd/fortran/projection.f90) keeps
the reference's `projection` interface and reaches HIP through the C ABI.
A Fortran driver fills the global arrays with the reference's bounds
(rhs_p(2:nxg,2:nyg-1,2:nzg)), calls compute_projection_step at RK stage 3 and
check_divergence, and the results are compared with the C oracle.  In its
device mode the driver copies the fields to device memory it allocates and
calls the device-resident entry points (compute_projection_step_device,
check_divergence_device), whose results must equal the host entry points'
bit for bit."""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import dfp
from oracle import ref_c
from _cases import fields, make_grid, rel_linf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = "/opt/rocm/lib/llvm/bin/flang"
DRIVER = os.path.join(ROOT, "tests", "fortran", "build", "driver")
SRC = [os.path.join(ROOT, "dopamine-fdm_amd", "fortran", "dfp_capi.f90"),
       os.path.join(ROOT, "tests", "fortran", "global.f90"),
       os.path.join(ROOT, "dopamine-fdm_amd", "fortran", "projection.f90"),
       os.path.join(ROOT, "tests", "fortran", "driver.f90")]


@pytest.mark.skipif(not os.path.exists(FLANG), reason="no Fortran compiler")
def test_dropin_compiles_and_links_against_the_c_abi():
    with tempfile.TemporaryDirectory() as d:
        objs = []
        for s in SRC:
            o = os.path.join(d, os.path.basename(s) + ".o")
            subprocess.run([FLANG, "-O1", f"-J{d}", "-c", s, "-o", o], check=True)
            objs.append(o)
        exe = os.path.join(d, "driver")
        subprocess.run([FLANG, *objs, "-o", exe, "-L" + os.path.dirname(dfp.LIB_PATH), "-ldfp",
                        "-L/opt/rocm/lib", "-lamdhip64"], check=True)
        assert os.path.exists(exe)


def _write_case(path, g, U, V, W, dt, beta=0.0, uc=0.0, Pb=None):
    if Pb is None:
        Pb = np.zeros((g.nyg - 2, g.nzg - 2), order="F")
    with open(path, "wb") as f:
        np.array([g.nxp, g.nym, g.nzp, getattr(g, "x_bc_type", 0), getattr(g, "y_bc_type", 1),
                  getattr(g, "z_bc_type", 0)], dtype=np.int32).tofile(f)
        np.array([g.dx, g.dz, dt, beta, uc]).tofile(f)
        for a in (g.x, g.y, g.z, g.yg, g.zg):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
        for a in (U, V, W, Pb):
            a.ravel(order="F").tofile(f)


def _read_out(path, g):
    raw = np.fromfile(path, dtype=np.float64)
    out, o = [], 0
    for shape in (g.shape_u, g.shape_v, g.shape_w, g.shape_rhs, g.shape_p):
        n = int(np.prod(shape))
        out.append(raw[o:o + n].reshape(shape, order="F"))
        o += n
    maxdiv = float(raw[o])
    phib = raw[o + 1:].reshape((g.nyg - 2, g.nzg - 2), order="F")
    return out, maxdiv, phib


# (grid dims, x_bc_type, outflow_backflow_beta); x_bc_type "py": the periodic box (y_bc_type = 0),
# "duct": the 4-wall duct (z_bc_type = 1, z faces stretched)
DROPIN_CASES = [((16, 12, 8), 0, 0.0), ((24, 17, 18), 0, 0.0), ((16, 10, 12), 1, 0.0), ((12, 9, 8), 1, 0.8),
                ((16, 13, 12), "py", 0.0), ((12, 10, 14), "duct", 0.0)]


@pytest.mark.gpu
@pytest.mark.parametrize("dims,xbc,beta", DROPIN_CASES, ids=lambda v: str(v))
def test_dropin_projection_matches_oracle(dims, xbc, beta, tmp_path):
    if not os.path.exists(DRIVER):
        raise RuntimeError(f"Fortran driver not built: {DRIVER} (run __graft_entry__.build())")
    if xbc == "py":
        g = dfp.periodic_y_grid(*dims)
    elif xbc == "duct":
        g = dfp.duct_grid(*dims, 1.8, 1.5)
    else:
        g = make_grid(*dims, 1.8, x_bc_type=xbc)
    U, V, W, P, rhs = fields(g, 21, 0.1)
    Pb = np.asfortranarray(np.random.default_rng(2).standard_normal((g.nyg - 2, g.nzg - 2)))
    uc = 0.9
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_case(fin, g, U, V, W, 1e-3, beta, uc, Pb)
    print(f"fortran driver {dims} x_bc_type {xbc} beta {beta}", flush=True)
    subprocess.run([DRIVER, str(fin), str(fout)], check=True, timeout=120)
    (Uf, Vf, Wf, rf, Pf), maxdiv, phibf = _read_out(fout, g)
    op = ref_c.CProjection(g)
    if beta > 0:
        op.set_outflow(beta, uc, Pb)
    op.compute_projection_step(U, V, W, rhs, P, 1e-3, 0.75)
    for lab, a, b in zip(("U", "V", "W", "phi", "P"), (Uf, Vf, Wf, rf, Pf), (U, V, W, rhs, P)):
        assert rel_linf(a, b) <= 1e-12, lab
    if beta > 0:
        assert rel_linf(phibf, op.phib) <= 1e-12
    # check_divergence on the driver's own corrected fields: the same stencil, bit-identical
    assert maxdiv == op.check_divergence(Uf, Vf, Wf)
    # and close to the oracle's value on its own fields (which differ at round-off)
    assert abs(maxdiv - op.check_divergence(U, V, W)) <= 1e-12 * max(1.0, abs(maxdiv))


@pytest.mark.gpu
@pytest.mark.parametrize("dims,xbc,beta", [DROPIN_CASES[0], DROPIN_CASES[3], DROPIN_CASES[4], DROPIN_CASES[5]],
                         ids=lambda v: str(v))
def test_dropin_device_entry_points_match_host_entry_points(dims, xbc, beta, tmp_path):
    if not os.path.exists(DRIVER):
        raise RuntimeError(f"Fortran driver not built: {DRIVER} (run __graft_entry__.build())")
    if xbc == "py":
        g = dfp.periodic_y_grid(*dims)
    elif xbc == "duct":
        g = dfp.duct_grid(*dims, 1.8, 1.5)
    else:
        g = make_grid(*dims, 1.8, x_bc_type=xbc)
    U, V, W, _, _ = fields(g, 23, 0.1)
    Pb = np.asfortranarray(np.random.default_rng(3).standard_normal((g.nyg - 2, g.nzg - 2)))
    fin = tmp_path / "in.bin"
    _write_case(fin, g, U, V, W, 1e-3, beta, 0.9, Pb)
    res = {}
    for mode in ("host", "device"):
        fout = tmp_path / f"out_{mode}.bin"
        print(f"fortran driver {dims} x_bc_type {xbc} beta {beta} {mode}", flush=True)
        subprocess.run([DRIVER, str(fin), str(fout), mode], check=True, timeout=120)
        res[mode] = np.fromfile(fout, dtype=np.float64)
    assert res["host"].shape == res["device"].shape
    assert np.array_equal(res["host"], res["device"])
