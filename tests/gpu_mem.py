"""Device copies of host arrays for the GPU tests, through the HIP runtime's
C API (ctypes): a reference result is uploaded once and compared on the
device with a library's state (Formura_Diff_Device), so that a test of many
kernel variants does not download the state for each one."""
from __future__ import annotations

import ctypes

import numpy as np

_hip = None


def _rt():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
    return _hip


class DeviceArray:
    """A device copy of a C-contiguous float64 or float32 array."""

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        if host.dtype not in (np.float64, np.float32):
            raise TypeError("float64 or float32 only")
        self.elem_size = host.dtype.itemsize
        self.nbytes = host.nbytes
        self.ptr = ctypes.c_void_p()
        rt = _rt()
        if rt.hipMalloc(ctypes.byref(self.ptr), self.nbytes) != 0:
            raise RuntimeError("hipMalloc failed")
        if rt.hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, 1) != 0:   # hipMemcpyHostToDevice
            raise RuntimeError("hipMemcpy failed")

    @property
    def address(self) -> int:
        return int(self.ptr.value)

    def free(self) -> None:
        if self.ptr:
            _rt().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def download(ptr: int, shape, dtype=np.float64) -> np.ndarray:
    """Copy `shape` elements of `dtype` from device address `ptr` to the host."""
    out = np.empty(shape, dtype=dtype)
    if _rt().hipMemcpy(out.ctypes.data, ctypes.c_void_p(ptr), out.nbytes, 2) != 0:   # hipMemcpyDeviceToHost
        raise RuntimeError("hipMemcpy failed")
    return out
