"""Device memory, streams and events through the HIP runtime libpintt.so uses.

torch ships its own HIP runtime build; two runtimes in one process do not
share a device, so the benchmark and tests drive device buffers, streams and
HIP events through /opt/rocm's libamdhip64 -- the very instance the engine
library already loaded -- via ctypes.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import load_library

H2D, D2H, D2D = 1, 2, 3

_rt = None


def runtime() -> ctypes.CDLL:
    global _rt
    if _rt is None:
        load_library()  # brings in libamdhip64.so.7; dlopen by soname reuses it
        rt = ctypes.CDLL("libamdhip64.so.7")
        vp = ctypes.c_void_p
        rt.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        rt.hipFree.argtypes = [vp]
        rt.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        rt.hipMemsetAsync.argtypes = [vp, ctypes.c_int, ctypes.c_size_t, vp]
        rt.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        rt.hipStreamDestroy.argtypes = [vp]
        rt.hipStreamSynchronize.argtypes = [vp]
        rt.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        rt.hipEventDestroy.argtypes = [vp]
        rt.hipEventRecord.argtypes = [vp, vp]
        rt.hipEventSynchronize.argtypes = [vp]
        rt.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        rt.hipGetErrorString.restype = ctypes.c_char_p
        rt.hipGetErrorString.argtypes = [ctypes.c_int]
        rt.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        rt.hipSetDevice.argtypes = [ctypes.c_int]
        _rt = rt
    return _rt


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: {runtime().hipGetErrorString(rc).decode()}")


def device_count() -> int:
    n = ctypes.c_int(0)
    rc = runtime().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


class DeviceBuffer:
    """A device allocation of `nbytes` (freed on close / garbage collection)."""

    def __init__(self, nbytes: int):
        self.nbytes = nbytes
        p = ctypes.c_void_p()
        check(runtime().hipMalloc(ctypes.byref(p), max(1, nbytes)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        buf = cls(a.nbytes)
        check(runtime().hipMemcpy(buf.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return buf

    def to_array(self, dtype=np.uint64, count: int | None = None) -> np.ndarray:
        itemsize = np.dtype(dtype).itemsize
        n = self.nbytes // itemsize if count is None else count
        out = np.empty(n, dtype=dtype)
        check(runtime().hipMemcpy(out.ctypes.data, self.ptr, n * itemsize, D2H), "hipMemcpy D2H")
        return out

    def close(self) -> None:
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        check(runtime().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self) -> None:
        check(runtime().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def close(self) -> None:
        if self.handle:
            runtime().hipStreamDestroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        check(runtime().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream: Stream | None = None) -> None:
        check(runtime().hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def synchronize(self) -> None:
        check(runtime().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, later: "Event") -> float:
        t = ctypes.c_float(0.0)
        check(runtime().hipEventElapsedTime(ctypes.byref(t), self.handle, later.handle), "hipEventElapsedTime")
        return float(t.value)

    def close(self) -> None:
        if self.handle:
            runtime().hipEventDestroy(self.handle)
            self.handle = None
