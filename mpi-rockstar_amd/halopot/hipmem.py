"""Device memory for Python callers, through the same HIP runtime that
libhalopot links (/opt/rocm's libamdhip64), so that one HSA runtime serves
the process.  (PyTorch wheels bundle their own HIP runtime; loading both in
one process leaves one of them without a device.)"""
from __future__ import annotations

import ctypes as C

import numpy as np

H2D, D2H, D2D = 1, 2, 3
_H = None


def hip() -> C.CDLL:
    global _H
    if _H is None:
        from . import lib

        lib()  # libhalopot first: its libamdhip64 is the process's runtime
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        h.hipGetErrorString.restype = C.c_char_p
        h.hipGetErrorString.argtypes = [C.c_int]
        _H = h
    return _H


def _ok(code: int, what: str) -> None:
    if code != 0:
        raise RuntimeError(f"{what} failed: {hip().hipGetErrorString(code).decode()}")


def device_count() -> int:
    n = C.c_int(0)
    if hip().hipGetDeviceCount(C.byref(n)) != 0:
        return 0
    return n.value


def synchronize() -> None:
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    """A non-blocking HIP stream at the default priority (it does not wait
    for the null stream); .handle is the hipStream_t for the C ABI."""

    def __init__(self):
        h = hip()
        h.hipStreamCreateWithFlags.argtypes = [C.POINTER(C.c_void_p), C.c_uint]
        h.hipStreamSynchronize.argtypes = [C.c_void_p]
        h.hipStreamDestroy.argtypes = [C.c_void_p]
        h.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        self.ptr = C.c_void_p()
        _ok(h.hipStreamCreateWithFlags(C.byref(self.ptr), 1), "hipStreamCreateWithFlags")

    @property
    def handle(self) -> int:
        return int(self.ptr.value or 0)

    def copy(self, dst: int, src: int, nbytes: int) -> None:
        """Device-to-device copy queued on this stream (not waited for)."""
        _ok(hip().hipMemcpyAsync(C.c_void_p(dst), C.c_void_p(src), nbytes, D2D, self.ptr),
            "hipMemcpyAsync")

    def synchronize(self) -> None:
        _ok(hip().hipStreamSynchronize(self.ptr), "hipStreamSynchronize")

    def destroy(self) -> None:
        if self.ptr.value:
            hip().hipStreamDestroy(self.ptr)
            self.ptr = C.c_void_p()


class DeviceArray:
    """A device allocation holding a copy of a numpy array's bytes."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(self.ptr), max(1, self.nbytes)), "hipMalloc")

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.upload(a)
        return d

    @classmethod
    def zeros(cls, nbytes: int) -> "DeviceArray":
        d = cls(nbytes)
        _ok(hip().hipMemset(d.ptr, 0, max(1, d.nbytes)), "hipMemset")
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _ok(hip().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D), "hipMemcpy H2D")

    def download(self, a: np.ndarray) -> np.ndarray:
        assert a.flags.c_contiguous and a.nbytes <= self.nbytes
        _ok(hip().hipMemcpy(a.ctypes.data_as(C.c_void_p), self.ptr, a.nbytes, D2H), "hipMemcpy D2H")
        return a

    @property
    def address(self) -> int:
        return int(self.ptr.value or 0)

    def free(self) -> None:
        if self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
