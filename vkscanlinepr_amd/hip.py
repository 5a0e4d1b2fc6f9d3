"""Device memory, streams and events of the HIP runtime that libspr.so itself is linked against
(ctypes), for callers of the device-pointer entry points (spr_scan_i32) and for timing work on
the stream it runs on.  Loading libspr.so first makes its libamdhip64 the one this module binds,
so pointers and streams made here are valid for the library whatever other runtime the process
also carries.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib

_hip = None
MEMCPY_H2D, MEMCPY_D2H, MEMCPY_D2D = 1, 2, 3
STREAM_NON_BLOCKING = 1


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        lib()
        H = C.CDLL("libamdhip64.so.7")
        vp = C.c_void_p
        H.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
        H.hipFree.argtypes = [vp]
        H.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
        H.hipMemset.argtypes = [vp, C.c_int, C.c_size_t]
        H.hipDeviceSynchronize.argtypes = []
        H.hipStreamCreateWithFlags.argtypes = [C.POINTER(vp), C.c_uint]
        H.hipStreamDestroy.argtypes = [vp]
        H.hipStreamSynchronize.argtypes = [vp]
        H.hipEventCreate.argtypes = [C.POINTER(vp)]
        H.hipEventDestroy.argtypes = [vp]
        H.hipEventRecord.argtypes = [vp, vp]
        H.hipEventSynchronize.argtypes = [vp]
        H.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), vp, vp]
        H.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        H.hipGetErrorString.argtypes = [C.c_int]
        H.hipGetErrorString.restype = C.c_char_p
        _hip = H
    return _hip


def _ok(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


def device_count() -> int:
    n = C.c_int(0)
    rc = hip().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


class DeviceBuffer:
    """nbytes of device memory (hipMalloc), freed by free() or when collected."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        _ok(hip().hipMemcpy(C.c_void_p(self.ptr + offset), a.ctypes.data_as(C.c_void_p), a.nbytes,
                            MEMCPY_H2D), "hipMemcpy")

    def download(self, dtype, count: int, offset: int = 0) -> np.ndarray:
        out = np.empty(int(count), dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        _ok(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr + offset), out.nbytes,
                            MEMCPY_D2H), "hipMemcpy")
        return out

    def fill_bytes(self, value: int) -> None:
        _ok(hip().hipMemset(C.c_void_p(self.ptr), int(value), self.nbytes), "hipMemset")
        _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free(self) -> None:
        if self.ptr:
            hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream at the default priority."""

    def __init__(self):
        s = C.c_void_p()
        _ok(hip().hipStreamCreateWithFlags(C.byref(s), STREAM_NON_BLOCKING), "hipStreamCreateWithFlags")
        self.handle = s.value

    def synchronize(self) -> None:
        _ok(hip().hipStreamSynchronize(C.c_void_p(self.handle)), "hipStreamSynchronize")

    def destroy(self) -> None:
        if self.handle:
            hip().hipStreamDestroy(C.c_void_p(self.handle))
            self.handle = 0


class Event:
    def __init__(self):
        e = C.c_void_p()
        _ok(hip().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream: Stream) -> None:
        _ok(hip().hipEventRecord(C.c_void_p(self.handle), C.c_void_p(stream.handle)), "hipEventRecord")

    def elapsed_ms(self, end: "Event") -> float:
        _ok(hip().hipEventSynchronize(C.c_void_p(end.handle)), "hipEventSynchronize")
        ms = C.c_float()
        _ok(hip().hipEventElapsedTime(C.byref(ms), C.c_void_p(self.handle), C.c_void_p(end.handle)),
            "hipEventElapsedTime")
        return float(ms.value)

    def destroy(self) -> None:
        if self.handle:
            hip().hipEventDestroy(C.c_void_p(self.handle))
            self.handle = 0
