"""Device buffers through the HIP runtime that libscb.so itself links (found in
this process's mappings), so callers share its device context without loading
a second runtime."""
import ctypes as C

import numpy as np

_H = None


def _hip():
    global _H
    if _H is None:
        from . import api
        api.lib()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line:
                    path = line.split()[-1]
                    break
        if path is None:
            raise RuntimeError("libscb.so is loaded but no HIP runtime is mapped")
        h = C.CDLL(path)
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        h.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        _H = h
    return _H


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceBuffer:
    """nbytes of device memory, freed with the object."""

    def __init__(self, nbytes):
        self.nbytes = max(1, int(nbytes))
        p = C.c_void_p()
        _ok(_hip().hipMalloc(C.byref(p), self.nbytes), "hipMalloc")
        self.ptr = p.value

    def __del__(self):
        if getattr(self, "ptr", None):
            _hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        if a.nbytes:
            _ok(_hip().hipMemcpy(C.c_void_p(d.ptr), C.c_void_p(a.ctypes.data), a.nbytes, 1), "hipMemcpy H2D")
        return d

    def zero(self):
        _ok(_hip().hipMemset(C.c_void_p(self.ptr), 0, self.nbytes), "hipMemset")
        return self

    def to_array(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            _ok(_hip().hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(self.ptr), out.nbytes, 2), "hipMemcpy D2H")
        return out


def device_free_bytes():
    """Free device memory of the current device (hipMemGetInfo)."""
    free, total = C.c_size_t(), C.c_size_t()
    _ok(_hip().hipMemGetInfo(C.byref(free), C.byref(total)), "hipMemGetInfo")
    return free.value


def synchronize():
    _ok(_hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
