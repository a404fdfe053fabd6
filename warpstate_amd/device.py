"""Device buffers on the HIP runtime the scan library itself links.

Callers of the WS_MEM_DEVICE entry points need input and output buffers in
device memory. They are allocated here through the very HIP runtime instance
that libwarpstate_amd.so uses (resolved by its soname once the library is
loaded), so one process never runs two HIP runtimes side by side.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _native as N

_H2D, _D2H, _D2D = 1, 2, 3
_hip = None


def _rt():
    global _hip
    if _hip is None:
        N.load()  # pulls in the runtime the library links against
        lib = ctypes.CDLL("libamdhip64.so.7")
        lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        lib.hipFree.argtypes = [ctypes.c_void_p]
        lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        lib.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        lib.hipSetDevice.argtypes = [ctypes.c_int]
        lib.hipGetErrorString.restype = ctypes.c_char_p
        lib.hipGetErrorString.argtypes = [ctypes.c_int]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipMemset", "hipSetDevice", "hipDeviceSynchronize"):
            getattr(lib, f).restype = ctypes.c_int
        _hip = lib
    return _hip


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = _rt().hipGetErrorString(rc).decode(errors="replace")
        raise RuntimeError(f"{what} failed: {msg} ({rc})")


def synchronize() -> None:
    _check(_rt().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceBuffer:
    """A plain device allocation; `ptr` is the device address."""

    def __init__(self, nbytes: int, device: int = 0):
        rt = _rt()
        _check(rt.hipSetDevice(device), "hipSetDevice")
        p = ctypes.c_void_p()
        _check(rt.hipMalloc(ctypes.byref(p), max(1, int(nbytes))), f"hipMalloc({nbytes})")
        self.ptr = int(p.value)
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, arr: np.ndarray, device: int = 0) -> "DeviceBuffer":
        a = np.ascontiguousarray(arr)
        b = cls(a.nbytes, device)
        if a.nbytes:
            _check(_rt().hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, _H2D), "hipMemcpy H2D")
        return b

    def upload(self, arr: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(arr)
        if offset + a.nbytes > self.nbytes:
            raise ValueError("upload past the end of the buffer")
        _check(_rt().hipMemcpy(self.ptr + offset, a.ctypes.data, a.nbytes, _H2D), "hipMemcpy H2D")

    def download(self, dtype=np.uint8, count: Optional[int] = None, offset: int = 0) -> np.ndarray:
        dt = np.dtype(dtype)
        n = (self.nbytes - offset) // dt.itemsize if count is None else int(count)
        if offset + n * dt.itemsize > self.nbytes:
            raise ValueError("download past the end of the buffer")
        out = np.empty(n, dtype=dt)
        if n:
            _check(_rt().hipMemcpy(out.ctypes.data, self.ptr + offset, n * dt.itemsize, _D2H), "hipMemcpy D2H")
        return out

    def copy_from(self, other: "DeviceBuffer", nbytes: Optional[int] = None) -> None:
        """Device-to-device copy (the runtime's own copy kernel)."""
        n = min(self.nbytes, other.nbytes) if nbytes is None else int(nbytes)
        _check(_rt().hipMemcpy(self.ptr, other.ptr, n, _D2D), "hipMemcpy D2D")

    def zero(self) -> None:
        _check(_rt().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def free(self) -> None:
        if getattr(self, "ptr", 0):
            _check(_rt().hipFree(self.ptr), "hipFree")
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
