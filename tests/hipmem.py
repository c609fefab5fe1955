"""Minimal device buffers for GPU tests through the HIP runtime (ctypes):
hipMalloc / hipMemcpy / hipFree, so tests of device-pointer entry points
do not depend on another framework's device management."""
import ctypes
import ctypes.util
import os

import numpy as np

_hip = None


def _loaded_runtime():
    """Path of the HIP runtime this process has already loaded (torch's own,
    when torch was imported first), or None."""
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                path = line.split()[-1]
                if "libamdhip64.so" in os.path.basename(path):
                    return path
    except OSError:
        pass
    return None


def hip():
    global _hip
    if _hip is None:
        path = _loaded_runtime()
        if path is None:
            path = "/opt/rocm/lib/libamdhip64.so"
            if not os.path.exists(path):
                path = ctypes.util.find_library("amdhip64") or "libamdhip64.so"
        h = ctypes.CDLL(path)
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipFree.argtypes = [ctypes.c_void_p]
        _hip = h
    return _hip


class DeviceArray:
    """A device copy of a host numpy array."""

    def __init__(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        self.dtype, self.size, self.nbytes = a.dtype, a.size, a.nbytes
        p = ctypes.c_void_p()
        rc = hip().hipMalloc(ctypes.byref(p), max(a.nbytes, 1))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")
        self.ptr = p.value
        rc = hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1)  # host to device
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed ({rc})")

    def get(self) -> np.ndarray:
        """Copy back to a new host array."""
        out = np.empty(self.size, dtype=self.dtype)
        rc = hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2)  # device to host
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed ({rc})")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None


def read_elements(ptr: int, dtype, indices) -> np.ndarray:
    """Copies the elements at `indices` of a device array of `dtype` back to
    the host (one small copy each)."""
    dt = np.dtype(dtype)
    out = np.empty(len(indices), dtype=dt)
    base = out.ctypes.data
    for k, i in enumerate(indices):
        rc = hip().hipMemcpy(base + k * dt.itemsize, ptr + int(i) * dt.itemsize, dt.itemsize, 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed ({rc})")
    return out


def read_array(ptr: int, dtype, n: int) -> np.ndarray:
    """Copies the first n elements of a device array of `dtype` back to the
    host (one copy)."""
    dt = np.dtype(dtype)
    out = np.empty(int(n), dtype=dt)
    if n > 0:
        rc = hip().hipMemcpy(out.ctypes.data, int(ptr), out.nbytes, 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed ({rc})")
    return out
