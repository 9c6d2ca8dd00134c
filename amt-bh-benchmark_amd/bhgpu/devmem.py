"""Device buffers for callers of the *_device entry points, on the HIP
runtime libbhgpu.so itself links (libamdhip64.so.7). A second HIP runtime
in the same process (e.g. the one bundled with a framework) would not see
the GPU once this one holds it, so nothing here goes through one."""
import ctypes

import numpy as np

from . import lib as bl

H2D, D2H, D2D = 1, 2, 3
_hip = None


def hip():
    global _hip
    if _hip is None:
        bl.load()  # loads libamdhip64.so.7 as libbhgpu's dependency
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
            f.restype = ctypes.c_int
        _hip = h
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


class DeviceBuffer:
    """nbytes of device memory; .ptr is its address."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    def upload(self, arr, offset=0):
        a = np.ascontiguousarray(arr)
        assert offset + a.nbytes <= self.nbytes
        _ok(hip().hipMemcpy(self.ptr + offset, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")

    def download(self, arr, offset=0):
        assert arr.flags["C_CONTIGUOUS"] and offset + arr.nbytes <= self.nbytes
        _ok(hip().hipMemcpy(arr.ctypes.data, self.ptr + offset, arr.nbytes, D2H), "hipMemcpy D2H")
        return arr

    def copy_from(self, other, nbytes=None):
        n = self.nbytes if nbytes is None else nbytes
        _ok(hip().hipMemcpy(self.ptr, other.ptr, n, D2D), "hipMemcpy D2D")

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
