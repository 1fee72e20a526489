"""Minimal ctypes access to the HIP runtime that libgo2.so itself links
(device buffers, copies, streams) for callers that hold data in device memory.

The same process must not also load a second HIP runtime (for example the one
bundled with a PyTorch wheel): two runtimes in one process do not share
devices.
"""
from __future__ import annotations

import ctypes

from . import lib

_hip = None
H2D, D2H, D2D = 1, 2, 3


def hip() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        lib()  # loads libamdhip64 as libgo2's dependency
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        h.hipGetErrorString.restype = ctypes.c_char_p
        h.hipGetErrorString.argtypes = [ctypes.c_int]
        _hip = h
    return _hip


def _ck(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


class DeviceBuffer:
    """Owned device allocation."""

    def __init__(self, nbytes: int):
        self.ptr = ctypes.c_void_p()
        self.nbytes = max(int(nbytes), 1)
        _ck(hip().hipMalloc(ctypes.byref(self.ptr), self.nbytes), "hipMalloc")

    @property
    def addr(self) -> int:
        return self.ptr.value

    def upload(self, arr) -> None:
        assert arr.nbytes <= self.nbytes
        _ck(hip().hipMemcpy(self.ptr, arr.ctypes.data, arr.nbytes, H2D), "hipMemcpy H2D")

    def download(self, arr) -> None:
        assert arr.nbytes <= self.nbytes
        _ck(hip().hipMemcpy(arr.ctypes.data, self.ptr, arr.nbytes, D2H), "hipMemcpy D2H")

    def free(self) -> None:
        if self.ptr and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        _ck(hip().hipStreamCreate(ctypes.byref(self.handle)), "hipStreamCreate")

    def synchronize(self) -> None:
        _ck(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self) -> None:
        if self.handle.value:
            hip().hipStreamDestroy(self.handle)
            self.handle = ctypes.c_void_p()
