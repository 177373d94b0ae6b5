"""Device memory and streams for the tests, from the HIP runtime the library links (the one
/opt/rocm ships; torch's wheel carries another runtime, whose pointers and streams this library's
runtime does not know)."""

from __future__ import annotations

import ctypes

import numpy as np

from autan_amd import _lib

_H2D, _D2H = 1, 2


def _hip():
    _lib.euler()  # the library first: its libamdhip64 is the one loaded
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    return hip


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: HIP error {rc}")


class DeviceArray:
    """n FP64 values in device memory."""

    def __init__(self, n: int) -> None:
        self.hip = _hip()
        self.n = n
        p = ctypes.c_void_p()
        check(self.hip.hipMalloc(ctypes.byref(p), 8 * n), "hipMalloc")
        self.ptr = p.value

    def copy_from(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        check(self.hip.hipMemcpy(self.ptr, a.ctypes.data, 8 * self.n, _H2D), "hipMemcpy")

    def copy_from_async(self, a: np.ndarray, stream: int) -> None:
        """From pinned-or-pageable host memory on `stream` (the host array must outlive the copy)."""
        assert a.dtype == np.float64 and a.flags.c_contiguous and a.size == self.n
        check(self.hip.hipMemcpyAsync(self.ptr, a.ctypes.data, 8 * self.n, _H2D, stream), "hipMemcpyAsync")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n)
        check(self.hip.hipMemcpy(out.ctypes.data, self.ptr, 8 * self.n, _D2H), "hipMemcpy")
        return out

    def free(self) -> None:
        if self.ptr:
            check(self.hip.hipFree(self.ptr), "hipFree")
            self.ptr = None


class Stream:
    def __init__(self) -> None:
        self.hip = _hip()
        s = ctypes.c_void_p()
        check(self.hip.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self) -> None:
        check(self.hip.hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self) -> None:
        if self.handle:
            check(self.hip.hipStreamDestroy(self.handle), "hipStreamDestroy")
            self.handle = None
