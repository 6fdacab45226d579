"""Device buffers and streams for tests, from the HIP runtime that
libpartc.so itself is linked against (so handles and pointers are valid in
the library).  libpartc.so is loaded first; opening libamdhip64.so.7 then
returns that same, already loaded runtime."""
from __future__ import annotations

import ctypes as C

import numpy as np

_hip = None
H2D, D2H = 1, 2


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        import partc_amd as P

        P.lib()
        L = C.CDLL("libamdhip64.so.7")
        vp = C.c_void_p
        L.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
        L.hipFree.argtypes = [vp]
        L.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
        L.hipStreamCreate.argtypes = [C.POINTER(vp)]
        L.hipStreamDestroy.argtypes = [vp]
        L.hipStreamSynchronize.argtypes = [vp]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreate", "hipStreamDestroy", "hipStreamSynchronize",
                  "hipDeviceSynchronize"):
            getattr(L, f).restype = C.c_int
        _hip = L
    return _hip


class DeviceBuffer:
    """nbytes of device memory (hipMalloc), freed by free()."""

    def __init__(self, nbytes: int):
        self.nbytes = max(int(nbytes), 1)
        self.ptr = C.c_void_p()
        rc = hip().hipMalloc(C.byref(self.ptr), self.nbytes)
        assert rc == 0, rc

    def to_numpy(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if count:
            rc = hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes, D2H)
            assert rc == 0, rc
        return out

    def free(self) -> None:
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()


def stream_create() -> int:
    s = C.c_void_p()
    rc = hip().hipStreamCreate(C.byref(s))
    assert rc == 0, rc
    return int(s.value)


def stream_destroy(s: int) -> None:
    hip().hipStreamDestroy(C.c_void_p(s))
