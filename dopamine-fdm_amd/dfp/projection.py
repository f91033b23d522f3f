"""Host-side mirror of the reference's ``projection`` module over libdfp.

``Plan`` wraps one ``dfp_plan`` (one grid on one device).  ``Projection``
keeps the reference's subroutine names (src/projection.f90) and works on a
state object holding the ``global`` module arrays, so a test reads like the
reference's own call sequence:

    st = ChannelState(grid);  st.U[...] = ...;  st.rk_step = 3
    proj = Projection(st)
    proj.compute_projection_step()          # projection.f90:21-34
    proj.check_divergence()                 # projection.f90:607-639

Arrays may be numpy arrays (host; copied to the device and back each call)
or torch CUDA tensors (device-resident; no copies).  Errors raise ``DfpError``
(the reference aborts through gpu_abort, poisson_gpu.f90:51-56).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import DFP_ASYNC, DFP_CHECK_FINITE, DFP_FIELDS_ON_DEVICE, DFP_FP64, DFP_SAVE_P, DfpGrid, DfpTuning, check, lib
from .grid import ChannelGrid
from .hipmem import DeviceArray

# src/initialization.f90:654-660 (diagonal of the RK3 table)
RK_COEF_DIAG = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


def _ptr(a, shape, name):
    """Address of a Fortran-ordered float64 array, plus whether it is on the device."""
    if a is None:
        return None, None
    if isinstance(a, DeviceArray):
        if a.shape != tuple(shape):
            raise ValueError(f"{name}: expected device array of shape {shape}, got {a.shape}")
        return a.ptr, True
    if _is_torch(a):
        import torch

        if a.dtype != torch.float64:
            raise ValueError(f"{name}: float64 required")
        # a Fortran array of shape (n1,n2,n3) is a C-contiguous (n3,n2,n1) tensor,
        # or its transposed view
        if tuple(a.shape) == tuple(reversed(shape)) and a.is_contiguous():
            pass
        elif tuple(a.shape) == tuple(shape) and a.permute(2, 1, 0).is_contiguous():
            pass
        else:
            raise ValueError(f"{name}: expected Fortran layout {shape}")
        return a.data_ptr(), bool(a.is_cuda)
    if not isinstance(a, np.ndarray) or a.dtype != np.float64:
        raise ValueError(f"{name}: float64 numpy array or torch tensor required")
    if a.shape != tuple(shape) or not a.flags.f_contiguous:
        raise ValueError(f"{name}: expected a Fortran-ordered array of shape {shape}, got {a.shape}")
    if not a.flags.writeable:
        raise ValueError(f"{name}: array is read-only")
    return a.ctypes.data, False


class Plan:
    """One ``dfp_plan``: operators, twiddles and device scratch of one grid."""

    def __init__(self, g: ChannelGrid, device: int = 0, precision: int = DFP_FP64, sweeps: int | None = None,
                 tuning: tuple | None = None):
        """tuning: (z_split, x_columns, z_rows[, z_fused]) launch shapes, 0 = the library's choice."""
        self.g = g
        self.precision = precision
        self.device = device
        self._arrays = [np.ascontiguousarray(a, dtype=np.float64) for a in (g.x, g.y, g.z, g.yg, g.zg)]
        dp = ctypes.POINTER(ctypes.c_double)
        x, y, z, yg, zg = (a.ctypes.data_as(dp) for a in self._arrays)
        self._grid = DfpGrid(g.nxp, g.nym, g.nzp, g.dx, g.dz, x, y, z, yg, zg, int(getattr(g, "x_bc_type", 0)),
                             int(getattr(g, "y_bc_type", 1)), int(getattr(g, "z_bc_type", 0)))
        self._outflow = None  # keeps the arrays registered with dfp_set_outflow alive
        h = ctypes.c_void_p()
        if tuning is None:
            check(lib().dfp_plan_create(ctypes.byref(self._grid), precision, device, ctypes.byref(h)), "dfp_plan_create")
        else:
            t = DfpTuning(*(int(v) for v in tuning))
            check(lib().dfp_plan_create_tuned(ctypes.byref(self._grid), precision, device, ctypes.byref(t),
                                              ctypes.byref(h)), "dfp_plan_create_tuned")
        self._h = h
        self.sweeps = 1 if sweeps is None else int(sweeps)
        if sweeps is not None:
            self.set_refine_sweeps(sweeps)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().dfp_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- helpers -------------------------------------------------------------
    def _flags(self, *devs, save_p=False, asynchronous=False):
        devs = [d for d in devs if d is not None]
        if any(devs) and not all(devs):
            raise ValueError("mixing host and device arrays in one call")
        f = DFP_FIELDS_ON_DEVICE if (devs and devs[0]) else 0
        if save_p:
            f |= DFP_SAVE_P
        if asynchronous and f & DFP_FIELDS_ON_DEVICE:
            f |= DFP_ASYNC
        return f

    def set_refine_sweeps(self, sweeps: int):
        """Refinement sweeps of an FP32+refine plan (0..8)."""
        check(lib().dfp_set_refine_sweeps(self._h, int(sweeps)), "dfp_set_refine_sweeps", self._h)
        self.sweeps = int(sweeps)

    @property
    def tuning(self) -> tuple:
        """(z_split, x_columns, z_rows, z_fused) the plan runs with (z_fused 2: one
        fused z kernel, 1: the split pair)."""
        t = DfpTuning()
        check(lib().dfp_plan_tuning(self._h, ctypes.byref(t)), "dfp_plan_tuning", self._h)
        return (t.z_split, t.x_columns, t.z_rows, t.z_fused)

    def set_stream(self, stream_handle: int | None):
        check(lib().dfp_set_stream(self._h, stream_handle), "dfp_set_stream", self._h)

    # --- entry points -----------------------------------------------------------
    def project(self, U, V, W, rhs_p, P=None, dt=0.0, c=0.0, save_p=False, asynchronous=False, check_finite=False):
        """check_finite: refuse u* holding NaN or |u| > 1e8 (the monitor's guard,
        monitor.f90:32-38) with DfpError code DFP_E_NONFINITE before writing anything."""
        g = self.g
        pu, du = _ptr(U, g.shape_u, "U")
        pv, dv = _ptr(V, g.shape_v, "V")
        pw, dw = _ptr(W, g.shape_w, "W")
        pr, dr = _ptr(rhs_p, g.shape_rhs, "rhs_p")
        pp, dpp = _ptr(P, g.shape_p, "P")
        f = self._flags(du, dv, dw, dr, dpp, save_p=save_p, asynchronous=asynchronous)
        if check_finite:
            f |= DFP_CHECK_FINITE
        check(lib().dfp_project(self._h, pu, pv, pw, pr, pp, float(dt), float(c), f), "dfp_project", self._h)

    def set_outflow(self, beta: float, outflow_Uc: float, Pb=None, phib=None):
        """Outlet backflow state of the inflow/outflow x (global.f90:188-197):
        Pb and phib are (nyg-2, nzg-2) arrays, host or device alike the fields of
        the calls that follow; the plan reads Pb and writes phib at each call."""
        g = self.g
        shape = (g.nyg - 2, g.nzg - 2)
        pb, dpb = _ptr(Pb, shape, "Pb")
        pf, dpf = _ptr(phib, shape, "phib")
        if Pb is not None and phib is not None and dpb != dpf:
            raise ValueError("Pb and phib must both be host or both be device arrays")
        check(lib().dfp_set_outflow(self._h, float(beta), float(outflow_Uc), pb, pf), "dfp_set_outflow", self._h)
        self._outflow = (Pb, phib)

    def outflow_backflow(self, U, V, W, rhs_p, dt, c):
        """set_outflow_backflow_pressure (projection.f90:65-96) on its own."""
        g = self.g
        pu, du = _ptr(U, g.shape_u, "U")
        pv, dv = _ptr(V, g.shape_v, "V")
        pw, dw = _ptr(W, g.shape_w, "W")
        pr, dr = _ptr(rhs_p, g.shape_rhs, "rhs_p")
        f = self._flags(du, dv, dw, dr)
        check(lib().dfp_outflow_backflow(self._h, pu, pv, pw, pr, float(dt), float(c), f), "dfp_outflow_backflow",
              self._h)

    # --- the step either side of the projection (channel, device arrays) -------
    def _step_arrays(self, stage, U, V, W, Uo, Vo, Wo, Fu, Fv, Fw, coef):
        g = self.g
        ptrs = []
        for a, sh, name in ((U, g.shape_u, "U"), (V, g.shape_v, "V"), (W, g.shape_w, "W"),
                            (Uo, g.shape_u, "Uo"), (Vo, g.shape_v, "Vo"), (Wo, g.shape_w, "Wo")):
            p, dev = _ptr(a, sh, name)
            if not dev:
                raise ValueError(f"{name} must be device-resident")
            ptrs.append(p)
        shapes = ((g.nx - 2, g.nyg - 2, g.nzg - 2), (g.nxg - 2, g.ny - 2, g.nzg - 2), (g.nxg - 2, g.nyg - 2, g.nz - 2))
        rhs = []
        for F, sh, name in zip((Fu, Fv, Fw), shapes, ("Fu", "Fv", "Fw")):
            if len(F) < stage:
                raise ValueError(f"{name}: {stage} right-hand sides needed")
            arr = (ctypes.c_void_p * 3)()
            for s_ in range(stage):
                p, dev = _ptr(F[s_], sh, name)
                if not dev:
                    raise ValueError(f"{name} must be device-resident")
                arr[s_] = p
            rhs.append(arr)
        c = (ctypes.c_double * 3)(*[float(v) for v in list(coef)[:stage]])
        return ptrs, rhs, c

    def _alpha_ptrs(self, alpha_x, alpha_y, alpha_z):
        g = self.g
        out = []
        for a, sh, name in ((alpha_x, (g.nx, 2, g.nzg), "alpha_x"), (alpha_y, (g.nxg, 2, g.nzg), "alpha_y"),
                            (alpha_z, (g.nxg, 2, g.nz), "alpha_z")):
            p, dev = _ptr(a, sh, name)
            if a is not None and not dev:
                raise ValueError(f"{name} must be device-resident")
            out.append(p)  # None (periodic y takes no alphas) passes a null pointer
        return out

    def rk_update(self, stage, U, V, W, Uo, Vo, Wo, Fu, Fv, Fw, coef, dt, asynchronous=False):
        """time_integration.f90:256-258 / :394-396 / :526-531 on device arrays;
        Fu, Fv, Fw: sequences of interior-shaped right-hand sides, coef =
        rk_coef(stage, 1:stage)."""
        ptrs, rhs, c = self._step_arrays(stage, U, V, W, Uo, Vo, Wo, Fu, Fv, Fw, coef)
        f = DFP_FIELDS_ON_DEVICE | (DFP_ASYNC if asynchronous else 0)
        check(lib().dfp_rk_update(self._h, int(stage), *ptrs, *rhs, c, float(dt), f), "dfp_rk_update", self._h)

    def apply_bc(self, U, V, W, alpha_x=None, alpha_y=None, alpha_z=None, asynchronous=False):
        """apply_boundary_conditions (boundary_conditions.f90:186-260): the
        channel's Robin walls (alphas required) or the periodic y wrap."""
        g = self.g
        pu, du = _ptr(U, g.shape_u, "U")
        pv, dv = _ptr(V, g.shape_v, "V")
        pw, dw = _ptr(W, g.shape_w, "W")
        f = self._flags(du, dv, dw, asynchronous=asynchronous)
        check(lib().dfp_apply_bc(self._h, pu, pv, pw, *self._alpha_ptrs(alpha_x, alpha_y, alpha_z), f),
              "dfp_apply_bc", self._h)

    def substep(self, stage, U, V, W, Uo, Vo, Wo, Fu, Fv, Fw, coef, alpha_x, alpha_y, alpha_z, rhs_p, P=None,
                dt=0.0, save_p=False, asynchronous=False):
        """One RK3 substage around the projection (time_integration.f90:256-304)."""
        g = self.g
        ptrs, rhs, c = self._step_arrays(stage, U, V, W, Uo, Vo, Wo, Fu, Fv, Fw, coef)
        pr, dr = _ptr(rhs_p, g.shape_rhs, "rhs_p")
        pp, dpp = _ptr(P, g.shape_p, "P")
        if not dr or (P is not None and not dpp):
            raise ValueError("rhs_p and P must be device-resident")
        f = DFP_FIELDS_ON_DEVICE | (DFP_SAVE_P if save_p else 0) | (DFP_ASYNC if asynchronous else 0)
        check(lib().dfp_substep(self._h, int(stage), *ptrs, *rhs, c, *self._alpha_ptrs(alpha_x, alpha_y, alpha_z),
                                pr, pp, float(dt), f), "dfp_substep", self._h)

    def divergence(self, U, V, W, rhs_p):
        g = self.g
        pu, du = _ptr(U, g.shape_u, "U")
        pv, dv = _ptr(V, g.shape_v, "V")
        pw, dw = _ptr(W, g.shape_w, "W")
        pr, dr = _ptr(rhs_p, g.shape_rhs, "rhs_p")
        f = self._flags(du, dv, dw, dr)
        check(lib().dfp_divergence(self._h, pu, pv, pw, pr, f), "dfp_divergence", self._h)

    def solve(self, rhs_p, P=None, dt=0.0, c=0.0, save_p=False, asynchronous=False):
        g = self.g
        pr, dr = _ptr(rhs_p, g.shape_rhs, "rhs_p")
        pp, dpp = _ptr(P, g.shape_p, "P")
        f = self._flags(dr, dpp, save_p=save_p, asynchronous=asynchronous)
        check(lib().dfp_solve(self._h, pr, pp, float(dt), float(c), f), "dfp_solve", self._h)

    def correct(self, U, V, W, rhs_p):
        g = self.g
        pu, du = _ptr(U, g.shape_u, "U")
        pv, dv = _ptr(V, g.shape_v, "V")
        pw, dw = _ptr(W, g.shape_w, "W")
        pr, dr = _ptr(rhs_p, g.shape_rhs, "rhs_p")
        f = self._flags(du, dv, dw, dr)
        check(lib().dfp_correct(self._h, pu, pv, pw, pr, f), "dfp_correct", self._h)

    def max_divergence(self, U, V, W) -> float:
        g = self.g
        pu, du = _ptr(U, g.shape_u, "U")
        pv, dv = _ptr(V, g.shape_v, "V")
        pw, dw = _ptr(W, g.shape_w, "W")
        f = self._flags(du, dv, dw)
        out = ctypes.c_double()
        check(lib().dfp_max_divergence(self._h, pu, pv, pw, ctypes.byref(out), f), "dfp_max_divergence", self._h)
        return out.value

    def synth(self, U, V, W, seed: int, sigma: float, asynchronous: bool = False):
        """Device-side synthetic fields (device arrays in Fortran layout), on the plan's stream."""
        g = self.g
        pu, du = _ptr(U, g.shape_u, "U")
        pv, dv = _ptr(V, g.shape_v, "V")
        pw, dw = _ptr(W, g.shape_w, "W")
        if not (du and dv and dw):
            raise ValueError("synth needs device arrays")
        check(lib().dfp_synth_fields_ex(self._h, pu, pv, pw, ctypes.c_uint64(seed & (2 ** 64 - 1)), float(sigma),
                                        DFP_ASYNC if asynchronous else 0), "dfp_synth_fields_ex", self._h)

    # --- stage timing ---------------------------------------------------------------
    def profile(self, enable: bool):
        check(lib().dfp_profile_enable(self._h, 1 if enable else 0), "dfp_profile_enable", self._h)

    def profile_reset(self):
        check(lib().dfp_profile_reset(self._h), "dfp_profile_reset", self._h)

    def profile_read(self) -> dict:
        out = {}
        i = 0
        while True:
            name = ctypes.c_char_p()
            n = ctypes.c_int64()
            ms = ctypes.c_double()
            rc = lib().dfp_profile_read(self._h, i, ctypes.byref(name), ctypes.byref(n), ctypes.byref(ms))
            if rc != 0:
                break
            out[name.value.decode()] = (int(n.value), float(ms.value))
            i += 1
        return out


@dataclass
class ChannelState:
    """The ``global``-module data the projection reads and writes
    (global.f90:139-146, :177-178, :293-294): fields, dt and the RK stage."""

    g: ChannelGrid
    U: np.ndarray = field(init=False)
    V: np.ndarray = field(init=False)
    W: np.ndarray = field(init=False)
    P: np.ndarray = field(init=False)
    rhs_p: np.ndarray = field(init=False)
    dt: float = 1e-3
    rk_step: int = 1
    # outlet backflow stabilisation of the inflow/outflow x (global.f90:188-197)
    outflow_backflow_beta: float = 0.0
    outflow_Uc: float = 0.0
    Pb: np.ndarray = field(init=False)
    phib: np.ndarray = field(init=False)

    def __post_init__(self):
        g = self.g
        self.U = np.zeros(g.shape_u, order="F")
        self.V = np.zeros(g.shape_v, order="F")
        self.W = np.zeros(g.shape_w, order="F")
        self.P = np.zeros(g.shape_p, order="F")
        self.rhs_p = np.zeros(g.shape_rhs, order="F")
        self.Pb = np.zeros((g.nyg - 2, g.nzg - 2), order="F")
        self.phib = np.zeros((g.nyg - 2, g.nzg - 2), order="F")


class Projection:
    """The reference module's subroutines (src/projection.f90) over a plan."""

    def __init__(self, state: ChannelState, device: int = 0):
        self.st = state
        self.plan = Plan(state.g, device)

    def _save_p(self, skip_p_save):
        # projection.f90:109-110, :362
        return self.st.rk_step == 3 and not bool(skip_p_save)

    def _sync_outflow(self):
        st = self.st
        if st.outflow_backflow_beta > 0.0:
            self.plan.set_outflow(st.outflow_backflow_beta, st.outflow_Uc, st.Pb, st.phib)
        else:
            self.plan.set_outflow(0.0, st.outflow_Uc)

    def compute_projection_step(self):
        """projection.f90:21-34 (the fused device path)."""
        st = self.st
        self._sync_outflow()
        # c = rk_coef(rk_step,rk_step): the backflow term's factor (projection.f90:82)
        # and, at the last stage, P = phi/(dt*rk_coef(3,3)) (projection.f90:362-377)
        self.plan.project(st.U, st.V, st.W, st.rhs_p, st.P, st.dt, RK_COEF_DIAG[st.rk_step - 1],
                          save_p=self._save_p(False))

    def set_outflow_backflow_pressure(self):
        """projection.f90:65-96"""
        st = self.st
        self._sync_outflow()
        self.plan.outflow_backflow(st.U, st.V, st.W, st.rhs_p, st.dt, RK_COEF_DIAG[st.rk_step - 1])

    def compute_pseudo_pressure_rhs(self):
        st = self.st
        self.plan.divergence(st.U, st.V, st.W, st.rhs_p)

    def solve_poisson_equation(self, skip_p_save=None):
        st = self.st
        self.plan.solve(st.rhs_p, st.P, st.dt, RK_COEF_DIAG[2], save_p=self._save_p(skip_p_save))

    def project_velocity(self):
        st = self.st
        self._sync_outflow()
        self.plan.correct(st.U, st.V, st.W, st.rhs_p)

    def check_divergence(self) -> float:
        st = self.st
        return self.plan.max_divergence(st.U, st.V, st.W)
