"""Device memory, streams and events for callers that keep particles in GPU
memory, taken from the HIP runtime the library itself links (ROCm's
libamdhip64): one runtime per process, shared with the kernels.

This is plumbing for the device-resident verbs (a Julia caller would pass
ROCArray pointers instead). A second HIP runtime in the same process (for
example the one bundled with a PyTorch wheel) does not see the GPU once this
one holds it, so tests and benchmarks use these helpers.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

H2D, D2H, D2D = 1, 2, 3

_hip = None


def hip():
    """The HIP runtime already loaded as the library's dependency."""
    global _hip
    if _hip is None:
        _lib.load()
        h = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        sig = {
            "hipMalloc": [ctypes.POINTER(vp), sz],
            "hipFree": [vp],
            "hipMemcpy": [vp, vp, sz, ctypes.c_int],
            "hipMemset": [vp, ctypes.c_int, sz],
            "hipDeviceSynchronize": [],
            "hipStreamCreate": [ctypes.POINTER(vp)],
            "hipStreamDestroy": [vp],
            "hipStreamSynchronize": [vp],
            "hipEventCreate": [ctypes.POINTER(vp)],
            "hipEventDestroy": [vp],
            "hipEventRecord": [vp, vp],
            "hipEventSynchronize": [vp],
            "hipEventElapsedTime": [ctypes.POINTER(ctypes.c_float), vp, vp],
        }
        for name, args in sig.items():
            f = getattr(h, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        h.hipGetErrorString.argtypes = [ctypes.c_int]
        h.hipGetErrorString.restype = ctypes.c_char_p
        _hip = h
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise _lib.GempicError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


class DeviceArray:
    """n float64 values in GPU memory (hipMalloc)."""

    def __init__(self, n: int):
        self.size = int(n)
        self.nbytes = 8 * max(self.size, 1)
        self.ptr = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(self.ptr), self.nbytes), "hipMalloc")

    @classmethod
    def from_numpy(cls, a) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        d = cls(a.size)
        d.copy_from(a)
        return d

    @classmethod
    def zeros(cls, n: int) -> "DeviceArray":
        d = cls(n)
        _ok(hip().hipMemset(d.ptr, 0, d.nbytes), "hipMemset")
        return d

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        if a.size != self.size:
            raise ValueError(f"size {a.size} != {self.size}")
        _ok(hip().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), 8 * a.size, H2D),
            "hipMemcpy")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.size)
        _ok(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, 8 * self.size, D2H),
            "hipMemcpy")
        return out

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p is not None and p.value:
            hip().hipFree(p)
            self.ptr = None


class Stream:
    """A HIP stream at the default priority."""

    def __init__(self):
        self.handle = ctypes.c_void_p()
        _ok(hip().hipStreamCreate(ctypes.byref(self.handle)), "hipStreamCreate")

    def synchronize(self):
        _ok(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            hip().hipStreamDestroy(h)
            self.handle = None


class Event:
    """A HIP event (timing enabled)."""

    def __init__(self):
        self.handle = ctypes.c_void_p()
        _ok(hip().hipEventCreate(ctypes.byref(self.handle)), "hipEventCreate")

    def record(self, stream: Stream | None = None):
        _ok(hip().hipEventRecord(self.handle, None if stream is None else stream.handle),
            "hipEventRecord")

    def synchronize(self):
        _ok(hip().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        _ok(hip().hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle),
            "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            hip().hipEventDestroy(h)
            self.handle = None


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
