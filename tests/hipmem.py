"""Device buffers for GPU tests through the HIP runtime the library itself is
linked against (ctypes), so that tests need no second runtime in the process."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so", mode=C.RTLD_GLOBAL)
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    return _hip


def free_bytes() -> int:
    """Free device memory of the current device (hipMemGetInfo)."""
    free, total = C.c_size_t(0), C.c_size_t(0)
    if hip().hipMemGetInfo(C.byref(free), C.byref(total)) != 0:
        raise RuntimeError("hipMemGetInfo failed")
    return free.value


H2D, D2H = 1, 2


class DeviceArray:
    def __init__(self, nbytes: int):
        p = C.c_void_p()
        rc = hip().hipMalloc(C.byref(p), nbytes)
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")
        self.ptr, self.nbytes = p.value, nbytes

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        if hip().hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, H2D) != 0:
            raise RuntimeError("hipMemcpy failed")
        return d

    def to_numpy(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        if hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) != 0:
            raise RuntimeError("hipMemcpy failed")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None
