"""ctypes binding of the oracle (oracle/liboracle.so) and of the one part of
the reference that builds here (oracle/_ref/libcxsparse.so).

Test infrastructure only: imported by tests/ and by bench.py's CPU-baseline
leg, never by the engine.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Optional

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
CXS = HERE / "_ref" / "libcxsparse.so"


class OrcTensor(C.Structure):
    _fields_ = [
        ("d", C.c_uint32),
        ("nk", C.c_uint32),
        ("N", C.c_uint64),
        ("hedges", C.POINTER(C.c_uint32)),
        ("vals", C.POINTER(C.c_double)),
        ("dims", C.POINTER(C.c_uint32)),
        ("dim_tc", C.POINTER(C.c_uint32)),
    ]


class OrcResult(C.Structure):
    _fields_ = [
        ("d", C.c_uint32),
        ("N", C.c_uint64),
        ("hedges", C.POINTER(C.c_uint32)),
        ("vals", C.POINTER(C.c_double)),
        ("flops", C.c_uint64),
        ("rows", C.c_uint64),
        ("rows_done", C.c_uint64),
        ("t_build_b", C.c_double),
        ("t_build_a", C.c_double),
        ("t_preprocess", C.c_double),
        ("t_contract", C.c_double),
    ]


_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise ImportError(f"{LIB} missing: build with `make -C oracle`")
        L = C.CDLL(str(LIB))
        T, R = C.POINTER(OrcTensor), C.POINTER(OrcResult)
        L.orc_tc_fksCuckoo.argtypes = [T, T, C.c_uint64, R]
        L.orc_contract_meta.argtypes = [T, T, C.c_uint64, R]
        L.orc_matricize_spGEMM.argtypes = [T, T, R]
        L.orc_flops.argtypes = [T, T, C.POINTER(C.c_uint64)]
        L.orc_result_free.argtypes = [R]
        L.orc_result_free.restype = None
        L.orc_seed.argtypes = [C.c_uint64]
        L.orc_seed.restype = None
        L.orc_set_plant.argtypes = [C.c_int]
        L.orc_set_plant.restype = None
        L.orc_fastmod_M.argtypes = [C.c_uint32]
        L.orc_fastmod_M.restype = C.c_uint64
        L.orc_fastmod_u32.argtypes = [C.c_uint32, C.c_uint64, C.c_uint32]
        L.orc_fastmod_u32.restype = C.c_uint32
        _lib = L
    return _lib


class _Keep:
    """Keeps numpy arrays alive while an OrcTensor points at them."""

    def __init__(self, t):
        self.h = np.ascontiguousarray(t.hedges_array, dtype=np.uint32)
        self.v = np.ascontiguousarray(t.val_array, dtype=np.float64)
        self.dims = np.ascontiguousarray(t.dimensions_array, dtype=np.uint32)
        self.perm = np.ascontiguousarray(t.dim_tc, dtype=np.uint32)
        self.s = OrcTensor(self.h.shape[1], int(t.num_key_dimensions), self.h.shape[0],
                           self.h.ctypes.data_as(C.POINTER(C.c_uint32)), self.v.ctypes.data_as(C.POINTER(C.c_double)),
                           self.dims.ctypes.data_as(C.POINTER(C.c_uint32)),
                           self.perm.ctypes.data_as(C.POINTER(C.c_uint32)))


class OracleError(RuntimeError):
    def __init__(self, code):
        self.code = code
        super().__init__(f"oracle error {code}")


def _collect(res: OrcResult):
    n, d = int(res.N), int(res.d)
    if n:
        h = np.ctypeslib.as_array(res.hedges, shape=(n * d,)).reshape(n, d).copy()
        v = np.ctypeslib.as_array(res.vals, shape=(n,)).copy()
    else:
        h = np.zeros((0, d), dtype=np.uint32)
        v = np.zeros(0, dtype=np.float64)
    info = {k: getattr(res, k) for k in ("flops", "rows", "rows_done", "t_build_b", "t_build_a", "t_preprocess",
                                         "t_contract")}
    lib().orc_result_free(C.byref(res))
    return h, v, info


def tc_fksCuckoo(A, B, max_rows: int = 0, seed: int = 0x1234ABCD):
    """Oracle contraction (A free-first, B contracted-first), in the
    reference's emission order.  Returns (hedges, vals, info)."""
    L = lib()
    L.orc_seed(seed)
    ka, kb = _Keep(A), _Keep(B)
    res = OrcResult()
    rc = L.orc_tc_fksCuckoo(C.byref(ka.s), C.byref(kb.s), max_rows, C.byref(res))
    if rc != 0:
        L.orc_result_free(C.byref(res))
        raise OracleError(rc)
    return _collect(res)


def contract_meta(A, B, max_rows: int = 0, seed: int = 0x1234ABCD):
    """The driver path (main.cpp:68-118): both tensors .meta-style
    (contracted modes first); A_copy is formed free-first inside."""
    L = lib()
    L.orc_seed(seed)
    ka, kb = _Keep(A), _Keep(B)
    res = OrcResult()
    rc = L.orc_contract_meta(C.byref(ka.s), C.byref(kb.s), max_rows, C.byref(res))
    if rc != 0:
        L.orc_result_free(C.byref(res))
        raise OracleError(rc)
    return _collect(res)


def matricize_spGEMM(A, B):
    """Oracle of the matricized path (both contracted-first)."""
    L = lib()
    ka, kb = _Keep(A), _Keep(B)
    res = OrcResult()
    rc = L.orc_matricize_spGEMM(C.byref(ka.s), C.byref(kb.s), C.byref(res))
    if rc != 0:
        L.orc_result_free(C.byref(res))
        raise OracleError(rc)
    return _collect(res)


def flops(A, B) -> int:
    """F = sum_k |A_k| |B_k| over contracted keys (both contracted-first)."""
    ka, kb = _Keep(A), _Keep(B)
    f = C.c_uint64()
    rc = lib().orc_flops(C.byref(ka.s), C.byref(kb.s), C.byref(f))
    if rc != 0:
        raise OracleError(rc)
    return int(f.value)


def fastmod_u32(a: int, d: int) -> int:
    L = lib()
    return int(L.orc_fastmod_u32(a, L.orc_fastmod_M(d), d))


def sort_coo(h: np.ndarray, v: np.ndarray):
    if h.shape[0] == 0:
        return h, v
    order = np.lexsort(h.T[::-1])
    return h[order], v[order]


# -------------------------------------------------------------- CXSparse


class CsDi(C.Structure):
    """cs_di of extern/cxsparse/cs.h (int32 indices, double values)."""

    _fields_ = [
        ("nzmax", C.c_int32),
        ("m", C.c_int32),
        ("n", C.c_int32),
        ("p", C.POINTER(C.c_int32)),
        ("i", C.POINTER(C.c_int32)),
        ("x", C.POINTER(C.c_double)),
        ("nz", C.c_int32),
    ]


def cxsparse_available() -> bool:
    return CXS.exists()


def cxsparse_multiply(Ap, Ai, Ax, m, k, Bp, Bi, Bx, n):
    """C = A*B with the reference's CXSparse cs_multiply (cs_multiply.c:6)
    on CSC inputs; returns CSC (p, i, x) of C."""
    L = C.CDLL(str(CXS))
    L.cs_di_multiply.argtypes = [C.POINTER(CsDi), C.POINTER(CsDi)]
    L.cs_di_multiply.restype = C.POINTER(CsDi)
    L.cs_di_spfree.argtypes = [C.POINTER(CsDi)]
    L.cs_di_spfree.restype = C.c_void_p
    keep = []

    def mk(p, i, x, rows, cols):
        p = np.ascontiguousarray(p, dtype=np.int32)
        i = np.ascontiguousarray(i, dtype=np.int32)
        x = np.ascontiguousarray(x, dtype=np.float64)
        keep.extend([p, i, x])
        return CsDi(int(p[-1]), rows, cols, p.ctypes.data_as(C.POINTER(C.c_int32)),
                    i.ctypes.data_as(C.POINTER(C.c_int32)), x.ctypes.data_as(C.POINTER(C.c_double)), -1)

    a = mk(Ap, Ai, Ax, m, k)
    b = mk(Bp, Bi, Bx, k, n)
    cp = L.cs_di_multiply(C.byref(a), C.byref(b))
    if not cp:
        raise RuntimeError("cs_multiply returned NULL")
    c = cp.contents
    nnz = int(c.p[c.n])
    p = np.array([c.p[j] for j in range(c.n + 1)], dtype=np.int64)
    i = np.ctypeslib.as_array(c.i, shape=(max(nnz, 1),))[:nnz].copy()
    x = np.ctypeslib.as_array(c.x, shape=(max(nnz, 1),))[:nnz].copy()
    L.cs_di_spfree(cp)
    return p, i, x
