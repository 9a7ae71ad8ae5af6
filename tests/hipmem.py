"""Device buffers for tests, from the same HIP runtime that libdns_hip.so uses.

torch ships its own copy of the HIP runtime; two runtimes in one process do not share a device context,
so tests allocate through the libamdhip64 that the library itself loaded."""
import ctypes

import numpy as np

import dnship

_hip = None


def hip():
    global _hip
    if _hip is None:
        dnship.lib()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line:
                    path = line.split()[-1]
                    break
        L = ctypes.CDLL(path or "libamdhip64.so")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        _hip = L
    return _hip


class DeviceBuffer:
    H2D, D2H = 1, 2

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        rc = hip().hipMalloc(ctypes.byref(p), self.nbytes)
        assert rc == 0, "hipMalloc failed (%d)" % rc
        self.ptr = p.value
        assert hip().hipMemset(self.ptr, 0, self.nbytes) == 0

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        assert hip().hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, cls.H2D) == 0
        return b

    def to_array(self, shape, dtype):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
