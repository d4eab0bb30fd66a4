"""Minimal device-memory plumbing over the HIP runtime (ctypes), used to stage
device-resident input sequences for the throughput entry points."""
from __future__ import annotations

import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise ImportError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipSetDevice.argtypes = [C.c_int]
        _hip.hipGetErrorString.restype = C.c_char_p
        _hip.hipGetErrorString.argtypes = [C.c_int]
    return _hip


def _ok(e, what):
    if e != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(e).decode()}")


class DeviceArray:
    """A device copy of a numpy array (freed on close / garbage collection)."""

    def __init__(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        self.shape, self.dtype, self.nbytes = a.shape, a.dtype, a.nbytes
        p = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(p), max(a.nbytes, 1)), "hipMalloc")
        self.ptr = p.value
        if a.nbytes:
            _ok(hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")

    @classmethod
    def empty(cls, shape, dtype):
        return cls(np.zeros(shape, dtype))

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        synchronize()
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2), "hipMemcpy D2H")
        return out

    def copy_from(self, other: "DeviceArray"):
        _ok(hip().hipMemcpy(self.ptr, other.ptr, min(self.nbytes, other.nbytes), 3), "hipMemcpy D2D")

    def close(self):
        if getattr(self, "ptr", None):
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
