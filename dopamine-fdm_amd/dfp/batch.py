"""Batched projections over libdfp's batch executor (``dfp_batch_*``).

A ``Batch`` runs the projections of many plans - uneven grid sizes, FP64 and
FP32+refine mixed - with one persistent launch per stage (no counterpart in
the reference, which projects one grid per run).  Fields must be resident on
the device (``DeviceArray`` or CUDA tensors in the reference's Fortran
layout); results equal ``Plan.project`` on each case bit for bit.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

from ._lib import DFP_ASYNC, DFP_FIELDS_ON_DEVICE, DFP_SAVE_P, DFP_SOLVE_ONLY, DfpError, lib
from .projection import Plan, _ptr


@dataclass
class Case:
    """One projection of a batch: a plan and its device-resident fields."""
    plan: Plan
    U: object
    V: object
    W: object
    rhs: object
    P: object = None
    dt: float = 0.0
    c: float = 0.0
    save_p: bool = False


class Batch:
    def __init__(self, device: int = 0, max_cases: int = 64, graphs: bool = False, occupancy: tuple | None = None):
        h = ctypes.c_void_p()
        rc = lib().dfp_batch_create(device, max_cases, ctypes.byref(h))
        if rc:
            raise DfpError(rc, "dfp_batch_create")
        self._h = h
        self.device = device
        self.max_cases = max_cases
        if graphs:
            self._check(lib().dfp_batch_use_graphs(self._h, 1), "dfp_batch_use_graphs")
        if occupancy is not None:
            self.set_occupancy(*occupancy)

    def _check(self, rc, where):
        if rc:
            detail = lib().dfp_batch_last_error(self._h).decode() if self._h else ""
            raise DfpError(rc, where, detail)

    def set_occupancy(self, x_per_cu: int, z_per_cu: int, stencil_per_cu: int):
        """Workgroups per CU of the persistent x, z and stencil launches."""
        self._check(lib().dfp_batch_set_occupancy(self._h, int(x_per_cu), int(z_per_cu), int(stencil_per_cu)),
                    "dfp_batch_set_occupancy")

    def use_graphs(self, enable: bool):
        self._check(lib().dfp_batch_use_graphs(self._h, 1 if enable else 0), "dfp_batch_use_graphs")

    def set_parallel(self, aux_streams: int):
        """Independent launches of one stage run on up to 1 + aux_streams
        streams (0..3, default 3); 0 where several batches already overlap."""
        self._check(lib().dfp_batch_set_parallel(self._h, int(aux_streams)), "dfp_batch_set_parallel")

    def set_stream(self, stream_handle):
        self._check(lib().dfp_batch_set_stream(self._h, stream_handle), "dfp_batch_set_stream")

    def project(self, cases: list[Case], asynchronous: bool = False, solve_only: bool = False):
        """Project every case (one persistent launch per stage).  With
        solve_only, each case's rhs holds the right-hand side and receives phi
        (and P with save_p), as Plan.solve; U, V, W are not read (may be None)."""
        n = len(cases)
        if n > self.max_cases:
            raise ValueError(f"{n} cases exceed the batch capacity {self.max_cases}")
        VP = ctypes.c_void_p * max(n, 1)
        plans, U, V, W, R, P = VP(), VP(), VP(), VP(), VP(), VP()
        dt = (ctypes.c_double * max(n, 1))()
        cc = (ctypes.c_double * max(n, 1))()
        fl = (ctypes.c_int * max(n, 1))()
        for i, cs in enumerate(cases):
            g = cs.plan.g
            ptrs = []
            for a, shape, name in ((cs.U, g.shape_u, "U"), (cs.V, g.shape_v, "V"), (cs.W, g.shape_w, "W"),
                                   (cs.rhs, g.shape_rhs, "rhs_p")):
                if solve_only and name != "rhs_p" and a is None:
                    ptrs.append(None)
                    continue
                ptr, dev = _ptr(a, shape, name)
                if not dev:
                    raise ValueError(f"case {i}: {name} must be device-resident")
                ptrs.append(ptr)
            plans[i] = cs.plan.handle.value
            U[i], V[i], W[i], R[i] = ptrs
            if cs.save_p:
                pp, dev = _ptr(cs.P, g.shape_p, "P")
                if not dev:
                    raise ValueError(f"case {i}: P must be device-resident")
                P[i] = pp
            dt[i], cc[i] = float(cs.dt), float(cs.c)
            fl[i] = DFP_FIELDS_ON_DEVICE | (DFP_SAVE_P if cs.save_p else 0)
        bf = (DFP_ASYNC if asynchronous else 0) | (DFP_SOLVE_ONLY if solve_only else 0)
        self._check(lib().dfp_batch_project(self._h, n, plans, U, V, W, R, P, dt, cc, fl, bf), "dfp_batch_project")

    def close(self):
        if self._h:
            lib().dfp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
