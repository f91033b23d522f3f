"""Minimal HIP runtime access through ctypes: device buffers, copies, events.

Used by the tests and the benchmark to keep fields resident in HBM without
pulling in a tensor framework.  Buffers are Fortran-ordered float64 arrays.
"""
from __future__ import annotations

import ctypes
import ctypes.util

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        _hip.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        _hip.hipFree.argtypes = [vp]
        _hip.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        _hip.hipMemset.argtypes = [vp, ctypes.c_int, sz]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        _hip.hipEventRecord.argtypes = [vp, vp]
        _hip.hipEventSynchronize.argtypes = [vp]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        _hip.hipEventDestroy.argtypes = [vp]
        _hip.hipGetErrorString.restype = ctypes.c_char_p
        _hip.hipGetErrorString.argtypes = [ctypes.c_int]
        _hip.hipMemGetInfo.argtypes = [ctypes.POINTER(sz), ctypes.POINTER(sz)]
        _hip.hipEventQuery.argtypes = [vp]
        _hip.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(vp), ctypes.c_uint]
        _hip.hipStreamDestroy.argtypes = [vp]
        _hip.hipStreamSynchronize.argtypes = [vp]
    return _hip


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()} ({rc})")


H2D, D2H, D2D = 1, 2, 3


def device_count() -> int:
    n = ctypes.c_int(0)
    rc = hip().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    _chk(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info():
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    _chk(hip().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return free.value, total.value


class DeviceArray:
    """A float64 device buffer with a Fortran shape."""

    def __init__(self, shape, zero=False):
        self.shape = tuple(int(s) for s in shape)
        self.size = int(np.prod(self.shape))
        self.nbytes = 8 * self.size
        p = ctypes.c_void_p()
        _chk(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 8)), "hipMalloc")
        self.ptr = p.value
        if zero:
            _chk(hip().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    @classmethod
    def from_host(cls, a: np.ndarray):
        d = cls(a.shape)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.asfortranarray(a, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError("shape mismatch")
        _chk(hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")

    def download(self) -> np.ndarray:
        a = np.empty(self.shape, order="F")
        _chk(hip().hipMemcpy(a.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return a

    def copy_from(self, other: "DeviceArray"):
        if other.shape != self.shape:
            raise ValueError("shape mismatch")
        _chk(hip().hipMemcpy(self.ptr, other.ptr, self.nbytes, D2D), "hipMemcpy D2D")

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        _chk(hip().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.e = e.value

    def record(self, stream=None):
        _chk(hip().hipEventRecord(self.e, stream), "hipEventRecord")

    def synchronize(self):
        _chk(hip().hipEventSynchronize(self.e), "hipEventSynchronize")

    def done(self) -> bool:
        """True once the work recorded before this event has completed."""
        rc = hip().hipEventQuery(self.e)
        if rc == 0:
            return True
        if rc == 600:  # hipErrorNotReady
            return False
        _chk(rc, "hipEventQuery")
        return False

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        _chk(hip().hipEventElapsedTime(ctypes.byref(ms), self.e, end.e), "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        try:
            hip().hipEventDestroy(self.e)
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream at the default priority."""

    def __init__(self):
        s = ctypes.c_void_p()
        _chk(hip().hipStreamCreateWithFlags(ctypes.byref(s), 1), "hipStreamCreateWithFlags")
        self.handle = s.value

    def synchronize(self):
        _chk(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def close(self):
        if self.handle:
            hip().hipStreamDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
