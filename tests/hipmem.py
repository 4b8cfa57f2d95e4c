"""Raw device buffers for tests of entry points that take device pointers
(hipMalloc / hipMemcpy / hipFree through ctypes)."""
import ctypes

import numpy as np

_hip = None
H2D, D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        for f in (_hip.hipMalloc, _hip.hipFree, _hip.hipMemcpy):
            f.restype = ctypes.c_int
    return _hip


class DeviceArray:
    """A device copy of a numpy array; .ptr is its device address."""

    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(p), max(self.host.nbytes, 1)) == 0, "hipMalloc"
        self.ptr = p
        self.put(self.host)

    def put(self, host):
        host = np.ascontiguousarray(host, dtype=self.host.dtype)
        assert host.nbytes == self.host.nbytes
        assert hip().hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, H2D) == 0, "hipMemcpy H2D"

    def get(self):
        out = np.empty_like(self.host)
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0, "hipMemcpy D2H"
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
