"""Device buffers for tests of the device-pointer entry points, taken from the
same HIP runtime libmqw.so links (/opt/rocm/lib/libamdhip64.so.7).  torch ships
its own HIP runtime; two runtimes in one process do not share devices, so the
GPU tests never initialise torch."""
import ctypes as C

import numpy as np

_hip = None
H2D, D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        import multistageqw_amd as M
        M.load()  # the runtime libmqw.so resolved
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DeviceArray:
    def __init__(self, host: np.ndarray):
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), self.host.nbytes) == 0
        assert hip().hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def numpy(self) -> np.ndarray:
        out = np.empty_like(self.host)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()
