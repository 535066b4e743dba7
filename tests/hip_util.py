"""Device buffers for tests through the HIP runtime that libsolidcloth.so
already loaded (ctypes on libamdhip64.so.7): plumbing for entry points that
take device pointers. Tests that use it are marked gpu."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        h = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipMemsetAsync.argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]
        h.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipStreamDestroy.argtypes = [C.c_void_p]
        h.hipStreamSynchronize.argtypes = [C.c_void_p]
        _hip = h
    return _hip


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


H2D, D2H = 1, 2


class DeviceArray:
    """A device allocation holding a float32 array of a given shape."""

    def __init__(self, shape, fill_bytes=None):
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 4
        self.ptr = C.c_void_p()
        check(hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 4)), "hipMalloc")
        if fill_bytes is not None:
            check(hip().hipMemset(self.ptr, fill_bytes, self.nbytes), "hipMemset")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        d = cls(a.shape)
        check(hip().hipMemcpy(d.ptr, a.ctypes.data, d.nbytes, H2D), "hipMemcpy")
        return d

    def numpy(self):
        out = np.empty(self.shape, dtype=np.float32)
        check(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy")
        return out

    def fill(self, byte, stream=None):
        """Byte fill, on `stream` (a Stream; None = the null stream)."""
        check(hip().hipMemsetAsync(self.ptr, byte, self.nbytes, stream.h if stream else None), "hipMemsetAsync")

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def free_memory():
    free, total = C.c_size_t(), C.c_size_t()
    check(hip().hipMemGetInfo(C.byref(free), C.byref(total)), "hipMemGetInfo")
    return free.value


class Stream:
    def __init__(self):
        self.h = C.c_void_p()
        check(hip().hipStreamCreate(C.byref(self.h)), "hipStreamCreate")

    def synchronize(self):
        check(hip().hipStreamSynchronize(self.h), "hipStreamSynchronize")

    def __del__(self):
        try:
            hip().hipStreamDestroy(self.h)
        except Exception:
            pass
