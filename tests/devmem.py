"""Caller-owned device buffers for the GPU tests (the role a Rust host's own
allocations play): hipMalloc / hipMemcpy / hipFree from the HIP runtime that
libhairmi.so itself links (libamdhip64.so.7). Plumbing only: every result
the tests check comes from libhairmi.so."""
from __future__ import annotations

import ctypes as C

import numpy as np

_hip = None
H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        from bevy_hair_amd import _native

        _native.lib()  # loads libamdhip64.so.7 as the library's dependency first
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def _check(err: int, what: str) -> None:
    if err != 0:
        raise RuntimeError(f"{what} failed with hipError {err}")


class DeviceArray:
    """A device allocation with the shape and dtype of a host array."""

    def __init__(self, host: np.ndarray):
        self.shape, self.dtype = host.shape, host.dtype
        self.nbytes = max(16, host.nbytes)
        p = C.c_void_p()
        _check(hip().hipMalloc(C.byref(p), self.nbytes), "hipMalloc")
        self.ptr = int(p.value)
        self.upload(host)

    def upload(self, host: np.ndarray) -> None:
        h = np.ascontiguousarray(host, dtype=self.dtype)
        assert h.shape == self.shape
        _check(hip().hipMemcpy(self.ptr, h.ctypes.data, h.nbytes, H2D), "hipMemcpy H2D")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _check(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self) -> None:
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = 0
