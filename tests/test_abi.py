"""The C ABI library: it loads, exports every function include/dfp.h
declares, and rejects what this build does not support before touching a
device (no compute calls: these tests run without a GPU)."""
from __future__ import annotations

import ctypes
import re
import subprocess

import numpy as np
import pytest

import dfp
from dfp import _lib


def _declared_functions():
    txt = open(_lib.HEADER_PATH).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*|dfp_plan \*)\s*(dfp_\w+)\s*\(", txt, re.M)))


def test_header_declares_the_boundary():
    names = _declared_functions()
    for n in ("dfp_plan_create", "dfp_plan_destroy", "dfp_project", "dfp_divergence", "dfp_solve",
              "dfp_correct", "dfp_max_divergence", "dfp_set_stream", "dfp_strerror"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = dfp.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (dfp_\w+)$", out, re.M))
    for n in _declared_functions():
        assert n in exported, n
        assert hasattr(L, n)


def test_strerror_codes():
    for code in (0, -1, -2, -3, -4, -5, -99):
        assert isinstance(dfp.lib().dfp_strerror(code), bytes)
    assert dfp.lib().dfp_strerror(-2) == b"unsupported configuration"


def _grid_struct(g, bc=(0, 1, 0)):
    dp = ctypes.POINTER(ctypes.c_double)
    keep = [np.ascontiguousarray(a) for a in (g.x, g.y, g.z, g.yg, g.zg)]
    ptrs = [a.ctypes.data_as(dp) for a in keep]
    return _lib.DfpGrid(g.nxp, g.nym, g.nzp, g.dx, g.dz, *ptrs, *bc), keep


@pytest.mark.parametrize("bc", [(2, 1, 0), (1, 0, 0), (0, 2, 0), (0, 1, 2), (1, 1, 1), (0, 0, 1)])
def test_other_boundary_conditions_are_rejected(bc):
    g = dfp.channel_grid(16, 8, 16, 1.5)
    s, keep = _grid_struct(g, bc)
    h = ctypes.c_void_p()
    assert dfp.lib().dfp_plan_create(ctypes.byref(s), 0, 0, ctypes.byref(h)) == -2
    assert not h.value


def test_unsupported_sizes_and_arguments_are_rejected():
    L = dfp.lib()
    h = ctypes.c_void_p()
    # FFT length with a factor 5 (nxp/2 = 5)
    g = dfp.channel_grid(10, 8, 16, 1.5)
    s, keep = _grid_struct(g)
    assert L.dfp_plan_create(ctypes.byref(s), 0, 0, ctypes.byref(h)) == -2
    g = dfp.channel_grid(16, 8, 20, 1.5)   # nzp = 20 = 4*5
    s, keep = _grid_struct(g)
    assert L.dfp_plan_create(ctypes.byref(s), 0, 0, ctypes.byref(h)) == -2
    # null grid / null output
    assert L.dfp_plan_create(None, 0, 0, ctypes.byref(h)) == -1
    assert L.dfp_plan_create(ctypes.byref(s), 0, 0, None) == -1
    # a null plan
    assert L.dfp_project(None, None, None, None, None, None, 0.0, 0.0, 0) == -1
    assert L.dfp_plan_destroy(None) == -1


def test_odd_nxp_rejected_by_host_grid_and_abi():
    with pytest.raises(ValueError):
        dfp.channel_grid(15, 8, 16, 1.5)


def test_host_grid_matches_reference_rules():
    g = dfp.namelist_grid(18, 17, 18, 3.0, 2.0, 1.5, 2.0)  # chan_les_small
    assert (g.nxp, g.nym, g.nzp) == (16, 16, 16)
    assert g.dx == pytest.approx(3.0 / 17, rel=1e-14)
    assert g.shape_u == (18, 18, 19) and g.shape_v == (19, 17, 19) and g.shape_w == (19, 18, 18)
    assert g.shape_rhs == (18, 16, 18)


def test_precision_argument_is_checked():
    L = dfp.lib()
    g = dfp.channel_grid(16, 8, 16, 1.5)
    s, keep = _grid_struct(g)
    h = ctypes.c_void_p()
    for bad in (-1, 2, 7):
        assert L.dfp_plan_create(ctypes.byref(s), bad, 0, ctypes.byref(h)) == -1
        assert not h.value
    assert L.dfp_set_refine_sweeps(None, 2) == -1


def test_flag_values_match_the_header():
    txt = open(_lib.HEADER_PATH).read()
    vals = {k: int(v) for k, v in re.findall(r"\b(DFP_[A-Z0-9_]+)\s*=\s*(\d+)", txt)}
    for name in ("DFP_FIELDS_ON_DEVICE", "DFP_SAVE_P", "DFP_ASYNC", "DFP_SOLVE_ONLY", "DFP_CHECK_FINITE", "DFP_FP64",
                 "DFP_FP32_REFINE"):
        assert getattr(dfp, name) == vals[name], name


def test_batch_project_rejects_a_null_batch_without_a_device():
    # argument checks come before any device call
    assert dfp.lib().dfp_batch_project(None, 0, None, None, None, None, None, None, None, None, None, 0) != 0
    assert dfp.lib().dfp_batch_project(None, 0, None, None, None, None, None, None, None, None, None,
                                       dfp.DFP_SOLVE_ONLY) != 0


def test_inflow_outflow_precision_and_outflow_arguments():
    """x inflow/outflow plans are FP64 only, refused before any device call;
    the outlet backflow state is checked before it is kept."""
    L = dfp.lib()
    g = dfp.channel_grid(16, 8, 16, 1.5, x_bc_type=1)
    s, keep = _grid_struct(g, (1, 1, 0))
    h = ctypes.c_void_p()
    assert L.dfp_plan_create(ctypes.byref(s), dfp.DFP_FP32_REFINE, 0, ctypes.byref(h)) == -2
    assert not h.value
    assert L.dfp_set_outflow(None, 0.5, 1.0, None, None) == -1
    assert L.dfp_outflow_backflow(None, None, None, None, None, 1e-3, 0.75, 0) == -1


def test_inflow_outflow_host_grid_rules():
    # initialization.f90:331-338: the inflow/outflow x keeps all nxm = nx-1 cells
    g = dfp.namelist_grid(17, 9, 10, 4.0, 2.0, 1.0, 1.5, x_bc_type=1)
    assert (g.nxp, g.nx, g.nxg) == (16, 17, 18)
    assert g.shape_u == (17, 10, 11) and g.shape_v == (18, 9, 11) and g.shape_rhs == (17, 8, 10)
    assert g.dx == pytest.approx(4.0 / 16, rel=1e-14)
    assert g.n_modes == 16 * 8 * 8


def test_periodic_y_rules_before_any_device_call():
    """y_bc_type = 0 needs a periodic x (initialization.f90:96-97), a uniform
    y grid (:128-129), nym - 1 = 2^a 3^b independent rows and FP64; each is
    refused before the device is touched."""
    L = dfp.lib()
    h = ctypes.c_void_p()
    g = dfp.periodic_y_grid(16, 9, 16)
    s, keep = _grid_struct(g, (0, 0, 0))
    assert L.dfp_plan_create(ctypes.byref(s), dfp.DFP_FP32_REFINE, 0, ctypes.byref(h)) == -2
    s, keep = _grid_struct(dfp.periodic_y_grid(16, 8, 16), (0, 0, 0))       # nym - 1 = 7
    assert L.dfp_plan_create(ctypes.byref(s), 0, 0, ctypes.byref(h)) == -2
    s, keep = _grid_struct(dfp.channel_grid(16, 9, 16, 1.5), (0, 0, 0))     # tanh-stretched y
    assert L.dfp_plan_create(ctypes.byref(s), 0, 0, ctypes.byref(h)) == -2
    s, keep = _grid_struct(dfp.channel_grid(16, 1, 16, 0.0, grid_type=1), (0, 0, 0))  # nym = 1
    assert L.dfp_plan_create(ctypes.byref(s), 0, 0, ctypes.byref(h)) == -2
    assert not h.value
    with pytest.raises(ValueError):
        dfp.ChannelGrid(16, 9, 16, 7.0, 2.0, 3.5, 1.5, 2, 0, 0)  # periodic y on a stretched grid
    with pytest.raises(ValueError):
        dfp.ChannelGrid(16, 9, 16, 7.0, 2.0, 3.5, 0.0, 1, 1, 0)  # with the inflow/outflow x
    # the last y face of the box is one period above the first (projection.f90:350-354)
    g = dfp.periodic_y_grid(8, 9, 6, Ly=2.0)
    assert g.y[-2] - g.y[0] == pytest.approx(2.0, rel=1e-14)


def test_duct_rules_before_any_device_call():
    """z walls are built as the 4-wall duct (periodic x, y walls,
    initialization.f90:98-100) in FP64; the rest is refused before the device
    is touched."""
    L = dfp.lib()
    h = ctypes.c_void_p()
    g = dfp.duct_grid(16, 9, 12, 1.5, 1.2)
    s, keep = _grid_struct(g, (0, 1, 1))
    assert L.dfp_plan_create(ctypes.byref(s), dfp.DFP_FP32_REFINE, 0, ctypes.byref(h)) == -2
    for bc in ((1, 1, 1), (0, 0, 1)):
        s, keep = _grid_struct(g, bc)
        assert L.dfp_plan_create(ctypes.byref(s), 0, 0, ctypes.byref(h)) == -2
    assert not h.value
    with pytest.raises(ValueError):
        dfp.ChannelGrid(16, 9, 12, 7.0, 2.0, 2.0, 1.5, 2, 1, 1, 1)   # inflow x with z walls
    # z walls keep every cell: nz = nzm + 1 faces, nzg = nzm + 2
    assert (g.nz, g.nzg) == (13, 14) and g.shape_w == (g.nxg, g.nyg, 13) and g.shape_rhs == (g.nxg - 1, 9, 13)
    assert g.z[0] == 0.0 and abs(g.z[-1] - 2.0) < 1e-15
