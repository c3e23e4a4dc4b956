"""Device arrays for tests of the device-pointer entry points, allocated
through the HIP runtime the library itself links (libamdhip64.so.7, already
loaded by libvxl_sph.so), so pointers are valid in both."""
import ctypes as C

import numpy as np

import vxl_sph as V

_hip = None


def hip():
    global _hip
    if _hip is None:
        V.lib()  # loads the runtime first
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    """A device copy of a numpy array (f32 / u32), freed on close()."""

    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), max(self.host.nbytes, 4)) == 0
        assert hip().hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    @property
    def addr(self):
        return self.ptr.value

    def get(self):
        out = np.empty_like(self.host)
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def close(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()
