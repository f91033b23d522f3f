"""ctypes front end of the C oracle (oracle/dfp_oracle.c).  Test infrastructure only."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libdfp_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)


class _Grid(ctypes.Structure):
    _fields_ = [
        ("nxp", ctypes.c_int), ("nym", ctypes.c_int), ("nzp", ctypes.c_int),
        ("dx", ctypes.c_double), ("dz", ctypes.c_double),
        ("x", _dp), ("y", _dp), ("z", _dp), ("yg", _dp), ("zg", _dp),
        ("x_bc_type", ctypes.c_int), ("y_bc_type", ctypes.c_int), ("z_bc_type", ctypes.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"C oracle not built: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.ora_plan_create.restype = vp
        L.ora_plan_create.argtypes = [ctypes.POINTER(_Grid), ctypes.c_int]
        L.ora_plan_destroy.argtypes = [vp]
        L.ora_operators.argtypes = [vp, vp, vp, vp, vp, vp]
        L.ora_divergence.argtypes = [vp, vp, vp, vp, vp]
        L.ora_solve.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.c_double]
        L.ora_solve.restype = ctypes.c_int
        L.ora_correct.argtypes = [vp, vp, vp, vp, vp]
        L.ora_project.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_double, ctypes.c_double]
        L.ora_project.restype = ctypes.c_int
        L.ora_max_divergence.argtypes = [vp, vp, vp, vp]
        L.ora_max_divergence.restype = ctypes.c_double
        L.ora_zgtsv.argtypes = [ctypes.c_int, vp, vp, vp, vp]
        L.ora_zgtsv.restype = ctypes.c_int
        L.ora_fft_many.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_long, ctypes.c_long, ctypes.c_int]
        L.ora_fft_many.restype = ctypes.c_int
        L.ora_threads.argtypes = [vp]
        L.ora_threads.restype = ctypes.c_int
        L.ora_channel_grid.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 4 + [ctypes.c_int] + [vp] * 8
        L.ora_channel_grid_x.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 4 + [ctypes.c_int] * 2 + [vp] * 8
        L.ora_set_outflow.argtypes = [vp, ctypes.c_double, ctypes.c_double, vp]
        L.ora_set_outflow.restype = ctypes.c_int
        L.ora_phib.argtypes = [vp, vp]
        L.ora_dct4_many.argtypes = [ctypes.c_int, ctypes.c_int, vp]
        L.ora_rk_update.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_double]
        L.ora_apply_bc.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _check_f64_fortran(a, shape, name):
    if a.dtype != np.float64 or not a.flags.f_contiguous or a.shape != tuple(shape):
        raise ValueError(f"{name} must be a Fortran-ordered float64 array of shape {shape}")


class CProjection:
    """The C oracle's projection on numpy arrays laid out as the reference's."""

    def __init__(self, g, nthreads: int = 0):
        self.g = g
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (g.x, g.y, g.z, g.yg, g.zg)]
        x, y, z, yg, zg = self._keep
        gd = _Grid(g.nxp, g.nym, g.nzp, g.dx, g.dz,
                   x.ctypes.data_as(_dp), y.ctypes.data_as(_dp), z.ctypes.data_as(_dp),
                   yg.ctypes.data_as(_dp), zg.ctypes.data_as(_dp), int(getattr(g, "x_bc_type", 0)),
                   int(getattr(g, "y_bc_type", 1)), int(getattr(g, "z_bc_type", 0)))
        self._p = lib().ora_plan_create(ctypes.byref(gd), int(nthreads))
        if not self._p:
            raise ValueError("ora_plan_create failed")

    def close(self):
        if self._p:
            lib().ora_plan_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_outflow(self, beta, outflow_Uc, Pb):
        """Outlet backflow state (x inflow/outflow): Pb has shape (nyg-2, nzg-2)."""
        g = self.g
        pb = None
        if Pb is not None:
            pb = np.asfortranarray(Pb, dtype=np.float64)
            if pb.shape != (g.nyg - 2, g.nzg - 2):
                raise ValueError("Pb must have shape (nyg-2, nzg-2)")
        self._pb = pb
        if lib().ora_set_outflow(self._p, float(beta), float(outflow_Uc), _ptr(pb)) != 0:
            raise ValueError("ora_set_outflow refused the backflow state")

    @property
    def phib(self):
        g = self.g
        out = np.zeros((g.nyg - 2, g.nzg - 2), order="F")
        lib().ora_phib(self._p, _ptr(out))
        return out

    @property
    def threads(self):
        return lib().ora_threads(self._p)

    def operators(self):
        g = self.g
        kxx = np.zeros(g.nxp); kzz = np.zeros(g.nzp)
        dl = np.zeros(max(g.nym - 1, 1)); d = np.zeros(g.nym); du = np.zeros(max(g.nym - 1, 1))
        lib().ora_operators(self._p, _ptr(kxx), _ptr(kzz), _ptr(dl), _ptr(d), _ptr(du))
        return kxx, kzz, dl[: g.nym - 1], d, du[: g.nym - 1]

    def _chk(self, U, V, W):
        g = self.g
        _check_f64_fortran(U, g.shape_u, "U")
        _check_f64_fortran(V, g.shape_v, "V")
        _check_f64_fortran(W, g.shape_w, "W")

    def compute_pseudo_pressure_rhs(self, U, V, W, rhs_p):
        self._chk(U, V, W)
        _check_f64_fortran(rhs_p, self.g.shape_rhs, "rhs_p")
        lib().ora_divergence(self._p, _ptr(U), _ptr(V), _ptr(W), _ptr(rhs_p))

    def solve_poisson_equation(self, rhs_p, P=None, dt=0.0, rk_coef33=0.0):
        _check_f64_fortran(rhs_p, self.g.shape_rhs, "rhs_p")
        if P is not None:
            _check_f64_fortran(P, self.g.shape_p, "P")
        rc = lib().ora_solve(self._p, _ptr(rhs_p), _ptr(P), float(dt or 0.0), float(rk_coef33 or 0.0))
        if rc != 0:
            raise FloatingPointError(f"oracle solve failed ({rc})")

    def project_velocity(self, U, V, W, rhs_p):
        self._chk(U, V, W)
        lib().ora_correct(self._p, _ptr(U), _ptr(V), _ptr(W), _ptr(rhs_p))

    def compute_projection_step(self, U, V, W, rhs_p, P=None, dt=0.0, rk_coef33=0.0):
        self._chk(U, V, W)
        _check_f64_fortran(rhs_p, self.g.shape_rhs, "rhs_p")
        if P is not None:
            _check_f64_fortran(P, self.g.shape_p, "P")
        rc = lib().ora_project(self._p, _ptr(U), _ptr(V), _ptr(W), _ptr(rhs_p), _ptr(P),
                               float(dt or 0.0), float(rk_coef33 or 0.0))
        if rc != 0:
            raise FloatingPointError(f"oracle projection failed ({rc})")

    def check_divergence(self, U, V, W):
        self._chk(U, V, W)
        return lib().ora_max_divergence(self._p, _ptr(U), _ptr(V), _ptr(W))

    def rk_update(self, stage, U, V, W, Uo, Vo, Wo, Fu, Fv, Fw, coef, dt):
        """time_integration.f90:256-258/:394-396/:526-531; Fu, Fv, Fw: lists of
        interior-shaped arrays (stage entries), coef = rk_coef(stage,1:stage)."""
        self._chk(U, V, W)
        self._chk(Uo, Vo, Wo)
        g = self.g
        shapes = ((g.nx - 2, g.nyg - 2, g.nzg - 2), (g.nxg - 2, g.ny - 2, g.nzg - 2), (g.nxg - 2, g.nyg - 2, g.nz - 2))
        arrs = []
        for F, sh, name in zip((Fu, Fv, Fw), shapes, ("Fu", "Fv", "Fw")):
            if len(F) < stage:
                raise ValueError(f"{name}: {stage} arrays needed")
            for a in F[:stage]:
                _check_f64_fortran(a, sh, name)
            arrs.append((ctypes.c_void_p * 3)(*[_ptr(a) for a in F[:stage]]))
        c = (ctypes.c_double * 3)(*[float(v) for v in coef[:stage]])
        lib().ora_rk_update(self._p, int(stage), _ptr(U), _ptr(V), _ptr(W), _ptr(Uo), _ptr(Vo), _ptr(Wo),
                            ctypes.cast(arrs[0], ctypes.c_void_p), ctypes.cast(arrs[1], ctypes.c_void_p),
                            ctypes.cast(arrs[2], ctypes.c_void_p), ctypes.cast(c, ctypes.c_void_p), float(dt))

    def apply_boundary_conditions(self, U, V, W, alpha_x, alpha_y, alpha_z):
        """boundary_conditions.f90:186-260 of the channel (or, y_bc_type = 0,
        the periodic box; its alphas may be None), one rank."""
        self._chk(U, V, W)
        g = self.g
        if getattr(g, "y_bc_type", 1) == 0 and alpha_x is None:
            lib().ora_apply_bc(self._p, _ptr(U), _ptr(V), _ptr(W), None, None, None)
            return
        _check_f64_fortran(alpha_x, (g.nx, 2, g.nzg), "alpha_x")
        _check_f64_fortran(alpha_y, (g.nxg, 2, g.nzg), "alpha_y")
        _check_f64_fortran(alpha_z, (g.nxg, 2, g.nz), "alpha_z")
        lib().ora_apply_bc(self._p, _ptr(U), _ptr(V), _ptr(W), _ptr(alpha_x), _ptr(alpha_y), _ptr(alpha_z))


def zgtsv(dl, d, du, b):
    """One system through the C restatement of zgtsv; returns (x, info)."""
    dl = np.array(dl, dtype=np.float64); d = np.array(d, dtype=np.float64); du = np.array(du, dtype=np.float64)
    x = np.array(b, dtype=np.complex128)
    info = lib().ora_zgtsv(d.size, _ptr(dl), _ptr(d), _ptr(du), _ptr(x))
    return x, info


def channel_grid_arrays(nxp, nym, nzp, Lx_i, Ly_i, Lz_i, alpha, grid_type=2, x_bc_type=0):
    nx, ny, nz = nxp + (1 if x_bc_type else 2), nym + 1, nzp + 2
    x = np.zeros(nx); y = np.zeros(ny); z = np.zeros(nz)
    xg = np.zeros(nx + 1); yg = np.zeros(ny + 1); zg = np.zeros(nz + 1)
    dx = ctypes.c_double(); dz = ctypes.c_double()
    lib().ora_channel_grid_x(nxp, nym, nzp, Lx_i, Ly_i, Lz_i, alpha, grid_type, int(x_bc_type),
                             _ptr(x), _ptr(y), _ptr(z), _ptr(xg), _ptr(yg), _ptr(zg),
                             ctypes.addressof(dx), ctypes.addressof(dz))
    return dict(x=x, y=y, z=z, xg=xg, yg=yg, zg=zg, dx=dx.value, dz=dz.value)


def dct4(a, axis=0):
    """DCT-IV (REDFT11 scaling) through the C restatement, along ``axis``."""
    b = np.array(np.moveaxis(np.asarray(a, dtype=np.float64), axis, -1), order="C", copy=True)
    n = b.shape[-1]
    lib().ora_dct4_many(n, b.size // n, _ptr(b))
    return np.moveaxis(b, -1, axis)
