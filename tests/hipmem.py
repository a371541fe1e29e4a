"""Device buffers for GPU tests through the HIP runtime's C API (ctypes).

Test plumbing only: uploads numpy arrays to device memory so entry points that
take a caller's device lattice can be called through the C ABI.
"""
import ctypes as C

import numpy as np

_hip = C.CDLL("libamdhip64.so")
_hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
_hip.hipFree.argtypes = [C.c_void_p]
_hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
_hip.hipDeviceSynchronize.argtypes = []
H2D, D2H = 1, 2


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed with HIP error %d" % (what, rc))


class DeviceArray:
    """A device copy of a numpy array (or `nbytes` uninitialised bytes)."""

    def __init__(self, host=None, nbytes=None):
        self.nbytes = int(host.nbytes if host is not None else nbytes)
        self.ptr = C.c_void_p()
        _ok(_hip.hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")
        if host is not None:
            self.upload(host)

    def upload(self, host):
        host = np.ascontiguousarray(host)
        assert host.nbytes <= self.nbytes
        _ok(_hip.hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, H2D), "hipMemcpy H2D")

    def download(self, dtype, count):
        out = np.empty(count, dtype)
        _ok(_hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
        _ok(_hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    @property
    def address(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            _hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        self.free()
