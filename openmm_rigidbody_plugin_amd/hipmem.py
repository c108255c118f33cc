"""Device buffers owned by the caller (the tests, the benchmark), standing in
for the arrays an OpenMM HIP context owns (posq, posqCorrection, velm, the
fixed-point force array) when they are handed to rbhip_adopt_arrays.

They come from the HIP runtime the library itself links (the same loaded
libamdhip64), so the library's kernels read and write them as a caller's
memory."""
import ctypes
import os

import numpy as np

_hip = None


def runtime():
    """The HIP runtime already loaded by librbhip.so."""
    global _hip
    if _hip is None:
        from . import rbhip
        rbhip.lib()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                p = line.split()[-1]
                if os.path.basename(p).startswith("libamdhip64.so"):
                    path = p
                    break
        if path is None:
            raise RuntimeError("the HIP runtime is not loaded")
        h = ctypes.CDLL(path)
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        for fn in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
            fn.restype = ctypes.c_int
        _hip = h
    return _hip


class DeviceArray:
    """A device copy of a numpy array's shape and dtype."""

    def __init__(self, like):
        self.shape, self.dtype = like.shape, like.dtype
        self.nbytes = like.nbytes
        self.ptr = ctypes.c_void_p()
        rc = runtime().hipMalloc(ctypes.byref(self.ptr), self.nbytes)
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")
        self.upload(like)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        h = runtime()
        assert h.hipDeviceSynchronize() == 0
        assert h.hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, 1) == 0   # host to device

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        h = runtime()
        assert h.hipDeviceSynchronize() == 0
        assert h.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        if self.ptr.value:
            runtime().hipDeviceSynchronize()
            runtime().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
