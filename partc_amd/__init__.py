"""partc_amd -- Python host interface of the MI355X sparse tensor contraction
engine (libpartc.so, C ABI in include/partc.h).

The names mirror the reference's C++ API (ssomesh/parTC):

    Tensor                 struct Tensor (tensor.hpp:19-174), numpy-backed
    Tensor.populate_meta_info   tensor.hpp:151-173 (.meta: 2 text lines, c, d mode ids)
    make_free_first        the A_copy construction of main.cpp:68-109
    get_sparse_tensor      sparse_tensor_generator.hpp:4-88 (B := A)
    buildHashTable         fksCuckoo.hpp:1044-1060 (+ reorder_tensor, fksCuckooTC.hpp:31-77)
    tc_fksCuckoo           fksCuckooTC.hpp:80-371
    matricize_spGEMM       matricizedspGEMM.hpp:172-603
    matricize_sort_spGEMM  matricizedspGEMM_sorting.hpp:194 (the reference exits
                           before computing it; here it returns the product)

Every compute call runs the HIP kernels of libpartc.so on the GPU; there is
no CPU fallback.  A missing library or device raises.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref
from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libpartc.so"
MAX_MODES = 8

PTC_OK = 0
ERRORS = {
    -1: "invalid argument",
    -2: "HIP runtime error",
    -3: "out of memory",
    -4: "tensor contraction not possible",
    -5: "range exceeded",
    -6: "unsupported",
    -7: "I/O error",
    -8: "no HIP device",
    -9: "internal consistency check failed",
}

METHOD_TWO_PASS, METHOD_LEARNED, METHOD_UPPER_BOUND = 0, 1, 2
VAL_F64, VAL_F32_TWOSUM, VAL_F32_PLAIN, VAL_F32_F64ACC = 0, 1, 2, 3
ACC_AUTO, ACC_DENSE, ACC_HASH, ACC_SPILL, ACC_WINDOW = 0, 1, 2, 3, 4
DIST_UNIFORM, DIST_ZIPF = 0, 1
GROUP_AUTO, GROUP_HASH, GROUP_SORT, GROUP_PARTITION = 0, 1, 2, 3


class PtcError(RuntimeError):
    def __init__(self, code: int, msg: str = ""):
        self.code = code
        super().__init__(f"{ERRORS.get(code, 'error')} ({code}){': ' + msg if msg else ''}")


class TensorView(C.Structure):
    _fields_ = [
        ("d", C.c_uint32),
        ("num_key_dimensions", C.c_uint32),
        ("N", C.c_uint64),
        ("hedges", C.POINTER(C.c_uint32)),
        ("dims", C.POINTER(C.c_uint32)),
        ("vals", C.POINTER(C.c_double)),
        ("dim_tc", C.POINTER(C.c_uint32)),
    ]


class Options(C.Structure):
    _fields_ = [
        ("method", C.c_int),
        ("value_type", C.c_int),
        ("acc_mode", C.c_int),
        ("dense_max_cols", C.c_uint32),
        ("lds_hash_max", C.c_uint32),
        ("chunk_flops", C.c_uint64),
        ("predict_scale", C.c_double),
        ("mlp", C.c_void_p),
        ("deterministic", C.c_int),
        ("block_threads", C.c_uint32),
    ]


class BatchStats(C.Structure):
    _fields_ = [
        ("jobs", C.c_uint32),
        ("bins", C.c_uint32),
        ("queue_items", C.c_uint32),
        ("runs", C.c_uint32),
        ("flops", C.c_uint64),
        ("nnz_out", C.c_uint64),
        ("graph_nodes", C.c_uint64),
        ("ms_build", C.c_double),
        ("ms_symbolic", C.c_double),
        ("ms_symbolic_pass", C.c_double),
        ("ms_plan", C.c_double),
        ("ms_numeric", C.c_double),
        ("ms_total", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneReport(C.Structure):
    _fields_ = [
        ("trials", C.c_uint32),
        ("dense_max_cols", C.c_uint32),
        ("lds_hash_max", C.c_uint32),
        ("chunk_flops", C.c_uint64),
        ("default_ms", C.c_double),
        ("best_ms", C.c_double),
        ("ms_wall", C.c_double),
        ("trial_dense", C.c_uint32 * 24),
        ("trial_hash", C.c_uint32 * 24),
        ("trial_chunk", C.c_uint64 * 24),
        ("trial_ms", C.c_double * 24),
    ]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("trial_")}
        n = min(self.trials, 24)
        d["trials_list"] = [(self.trial_dense[i], self.trial_hash[i], self.trial_chunk[i], round(self.trial_ms[i], 3))
                            for i in range(n)]
        return d


class HostTensor(C.Structure):
    _fields_ = [
        ("d", C.c_uint32),
        ("N", C.c_uint64),
        ("hedges", C.POINTER(C.c_uint32)),
        ("vals", C.POINTER(C.c_double)),
        ("dims", C.c_uint32 * 8),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("nnz_a", C.c_uint64),
        ("nnz_b", C.c_uint64),
        ("rows", C.c_uint64),
        ("flops", C.c_uint64),
        ("nnz_out", C.c_uint64),
        ("fallback_rows", C.c_uint64),
        ("heavy_rows", C.c_uint64),
        ("spill_rows", C.c_uint64),
        ("ms_build_b", C.c_double),
        ("ms_build_a", C.c_double),
        ("ms_preprocess", C.c_double),
        ("ms_symbolic", C.c_double),
        ("ms_numeric", C.c_double),
        ("ms_total", C.c_double),
        ("ms_kernel_symbolic", C.c_double),
        ("ms_kernel_numeric", C.c_double),
        ("kernel_launches_numeric", C.c_uint64),
        ("ms_build_out", C.c_double),
        ("out_capacity", C.c_uint64),
        ("device_peak_bytes", C.c_uint64),
        ("col_windows", C.c_uint64),
        ("small_rows", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """Load libpartc.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(str(LIB_PATH))
    vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    P32, P64, PD = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_double)
    sig = {
        "ptc_create": (i32, [C.POINTER(vp), i32]),
        "ptc_create_on_stream": (i32, [C.POINTER(vp), i32, vp]),
        "ptc_destroy": (None, [vp]),
        "ptc_last_error": (C.c_char_p, [vp]),
        "ptc_synchronize": (i32, [vp]),
        "ptc_set_timing": (i32, [vp, i32]),
        "ptc_stream": (vp, [vp]),
        "ptc_dtensor_upload": (i32, [vp, C.POINTER(TensorView), C.POINTER(vp)]),
        "ptc_dtensor_upload_f32": (i32, [vp, C.POINTER(TensorView), C.POINTER(C.c_float), C.POINTER(vp)]),
        "ptc_dtensor_wrap": (i32, [vp, u32, u32, u64, vp, vp, P32, P32, P32, C.POINTER(vp)]),
        "ptc_dtensor_set_keys": (i32, [vp, u32, P32]),
        "ptc_dtensor_info": (i32, [vp, P32, P64, P32]),
        "ptc_dtensor_device_ptrs": (i32, [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]),
        "ptc_dtensor_download": (i32, [vp, vp, P32, PD, P32]),
        "ptc_dtensor_download_f32": (i32, [vp, vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "ptc_dtensor_free": (None, [vp]),
        "ptc_meta_parse": (i32, [C.c_char_p, u32, P32, P32]),
        "ptc_hashtable_build": (i32, [vp, vp, C.POINTER(vp)]),
        "ptc_hashtable_build_ex": (i32, [vp, vp, i32, C.POINTER(vp)]),
        "ptc_hashtable_info": (i32, [vp, P64, P64]),
        "ptc_hashtable_probe": (i32, [vp, vp, u64, P32, P64, P64, P32, u64]),
        "ptc_hashtable_free": (None, [vp]),
        "ptc_options_default": (None, [C.POINTER(Options)]),
        "ptc_tc_fksCuckoo": (i32, [vp, vp, vp, vp, vp, C.POINTER(Options), C.POINTER(vp), C.POINTER(Stats)]),
        "ptc_tc_fksCuckoo_keyed": (i32, [vp, vp, vp, vp, vp, C.POINTER(Options), u32, P32, C.POINTER(vp),
                                         C.POINTER(vp), C.POINTER(i32), C.POINTER(Stats)]),
        "ptc_contract": (i32, [vp, vp, vp, C.POINTER(Options), C.POINTER(vp), C.POINTER(Stats)]),
        "ptc_tc_fksCuckoo_noalloc": (i32, [vp, vp, vp, vp, vp, C.POINTER(Options), vp, vp, u64, P64,
                                           C.POINTER(Stats)]),
        "ptc_matricize_spGEMM": (i32, [vp, vp, vp, C.POINTER(vp), C.POINTER(Stats)]),
        "ptc_matricize_sort_spGEMM": (i32, [vp, vp, vp, C.POINTER(vp), C.POINTER(Stats)]),
        "ptc_gen_coo": (i32, [u64, u32, P32, u64, i32, C.c_double, P32, PD]),
        "ptc_dtensor_set_values": (i32, [vp, vp, PD]),
        "ptc_read_tns": (i32, [C.c_char_p, C.POINTER(HostTensor)]),
        "ptc_read_mtx": (i32, [C.c_char_p, C.POINTER(HostTensor)]),
        "ptc_host_tensor_free": (None, [C.POINTER(HostTensor)]),
        "ptc_write_tns": (i32, [C.c_char_p, u32, u64, P32, PD]),
        "ptc_batch_plan": (i32, [vp, i32, C.POINTER(vp), C.POINTER(vp), C.POINTER(Options), C.POINTER(vp),
                                 C.POINTER(BatchStats)]),
        "ptc_batch_run": (i32, [vp, C.POINTER(BatchStats)]),
        "ptc_batch_output": (i32, [vp, i32, C.POINTER(vp)]),
        "ptc_batch_free": (None, [vp]),
        "ptc_batch_autotune": (i32, [vp, i32, C.POINTER(vp), C.POINTER(vp), C.POINTER(Options), C.POINTER(Options),
                                     C.POINTER(TuneReport)]),
        "ptc_contract_batched": (i32, [vp, i32, C.POINTER(vp), C.POINTER(vp), C.POINTER(Options), C.POINTER(vp),
                                       C.POINTER(BatchStats)]),
        "ptc_gen_job_shape": (i32, [u64, u32, u32, u64, u64, P32, P32, P64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p32(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _p64(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def _pd(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# --------------------------------------------------------------- host tensor


@dataclass
class Tensor:
    """Host COO tensor with the reference's fields (tensor.hpp:19-26)."""

    hedges_array: np.ndarray  # (N, d) uint32, 0-based
    val_array: np.ndarray  # (N,) float64
    dimensions_array: Optional[np.ndarray] = None  # (d,) uint32
    dim_tc: Optional[np.ndarray] = None  # (d,) uint32 permutation, key modes first
    num_key_dimensions: int = 0

    def __post_init__(self):
        h = np.ascontiguousarray(self.hedges_array, dtype=np.uint32)
        if h.ndim == 1:
            h = h.reshape(-1, 1)
        self.hedges_array = h
        self.val_array = np.ascontiguousarray(self.val_array, dtype=np.float64).reshape(-1)
        if self.val_array.shape[0] != h.shape[0]:
            raise ValueError("hedges and values disagree on N")
        if self.dimensions_array is None:
            self.dimensions_array = (h.max(axis=0) + 1 if h.shape[0] else np.zeros(h.shape[1])).astype(np.uint32)
        self.dimensions_array = np.ascontiguousarray(self.dimensions_array, dtype=np.uint32)
        if self.dim_tc is None:
            self.dim_tc = np.arange(h.shape[1], dtype=np.uint32)
        self.dim_tc = np.ascontiguousarray(self.dim_tc, dtype=np.uint32)

    @property
    def d(self) -> int:
        return int(self.hedges_array.shape[1])

    @property
    def N(self) -> int:
        return int(self.hedges_array.shape[0])

    def populate_meta_info(self, filename: str) -> None:
        """Read a .meta file: two free-text lines, c, then d mode ids whose
        first c are the contracted modes (tensor.hpp:151-173)."""
        c = C.c_uint32()
        perm = np.zeros(MAX_MODES, dtype=np.uint32)
        rc = lib().ptc_meta_parse(str(filename).encode(), self.d, C.byref(c), _p32(perm))
        if rc != PTC_OK:
            raise PtcError(rc, f"cannot parse {filename}")
        self.num_key_dimensions = int(c.value)
        self.dim_tc = perm[: self.d].copy()

    def write_tns(self, path) -> None:
        """FROSTT .tns text, 1-based coordinates (ptc_write_tns)."""
        rc = lib().ptc_write_tns(str(path).encode(), self.d, self.N, _p32(self.hedges_array), _pd(self.val_array))
        if rc != PTC_OK:
            raise PtcError(rc, f"cannot write {path}")

    def view(self) -> TensorView:
        return TensorView(self.d, int(self.num_key_dimensions), self.N, _p32(self.hedges_array),
                          _p32(self.dimensions_array), _pd(self.val_array), _p32(self.dim_tc))

    def sorted(self) -> "Tensor":
        """Copy with nonzeros in lexicographic coordinate order (the output
        order of a contraction is unspecified, fksCuckooTC.hpp:345)."""
        if self.N == 0:
            return self
        order = np.lexsort(self.hedges_array.T[::-1])
        return Tensor(self.hedges_array[order], self.val_array[order], self.dimensions_array, self.dim_tc,
                      self.num_key_dimensions)


def _read_host(fn, path) -> Tensor:
    t = HostTensor()
    rc = fn(str(path).encode(), C.byref(t))
    if rc != PTC_OK:
        raise PtcError(rc, f"cannot read {path}")
    try:
        n, d = int(t.N), int(t.d)
        h = np.ctypeslib.as_array(t.hedges, shape=(n * d,)).reshape(n, d).copy() if n else np.zeros((0, d), np.uint32)
        v = np.ctypeslib.as_array(t.vals, shape=(n,)).copy() if n else np.zeros(0)
        dims = np.array(t.dims[:d], dtype=np.uint32)
    finally:
        lib().ptc_host_tensor_free(C.byref(t))
    return Tensor(h, v, dims)


def read_tns(path) -> Tensor:
    """A FROSTT .tns file (readTensor, mtxTnsIO.cpp:31-63): 0-based
    coordinates, dims = largest label per mode."""
    return _read_host(lib().ptc_read_tns, path)


def read_mtx(path) -> Tensor:
    """A Matrix Market coordinate file (readMatrix, mtxTnsIO.cpp:5-29)."""
    return _read_host(lib().ptc_read_mtx, path)


def make_free_first(A: Tensor) -> Tensor:
    """A_copy of main.cpp:68-109: same data, key modes = A's free modes."""
    c = int(A.num_key_dimensions)
    perm = np.concatenate([A.dim_tc[c:], A.dim_tc[:c]]).astype(np.uint32)
    return Tensor(A.hedges_array, A.val_array, A.dimensions_array, perm, A.d - c)


def get_sparse_tensor(A: Tensor) -> Tensor:
    """B := A with A's key split (sparse_tensor_generator.hpp:4-88, factor 1)."""
    return Tensor(A.hedges_array.copy(), A.val_array.copy(), A.dimensions_array.copy(), A.dim_tc.copy(),
                  A.num_key_dimensions)


# ------------------------------------------------------------ device objects


def _check(rc: int, ctx: Optional["Context"] = None) -> None:
    if rc != PTC_OK:
        msg = ctx.last_error() if ctx is not None else ""
        raise PtcError(rc, msg)


_live_contexts: "weakref.WeakSet[Context]" = weakref.WeakSet()


def _close_live_contexts() -> None:
    """At interpreter exit, before the HIP runtime's own exit handlers: free
    every context's device objects and streams while the runtime is whole."""
    for c in list(_live_contexts):
        try:
            c.close()
        except Exception:
            pass


atexit.register(_close_live_contexts)


class Context:
    """One device and one HIP stream (ptc_create).  stream: a caller's
    hipStream_t (an int handle, e.g. torch.cuda.Stream().cuda_stream) that the
    context then orders all its work on (ptc_create_on_stream); the caller
    keeps it alive longer than the context."""

    def __init__(self, device: int = -1, stream: Optional[int] = None):
        self._h = C.c_void_p()
        # objects holding device memory of this context; freed before it
        self._children = weakref.WeakSet()
        if stream is None:
            rc = lib().ptc_create(C.byref(self._h), device)
        else:
            rc = lib().ptc_create_on_stream(C.byref(self._h), device, C.c_void_p(stream))
        if rc != PTC_OK:
            raise PtcError(rc, "ptc_create")
        _live_contexts.add(self)

    def adopt(self, obj):
        self._children.add(obj)
        return obj

    @property
    def handle(self):
        return self._h

    def last_error(self) -> str:
        s = lib().ptc_last_error(self._h)
        return s.decode() if s else ""

    def synchronize(self) -> None:
        _check(lib().ptc_synchronize(self._h), self)

    def set_timing(self, on: bool) -> None:
        """Phase timing by HIP events (ptc_set_timing): the ms_* statistics
        of builds and contractions; off by default, since every event idles
        the GPU for a few microseconds between kernels."""
        _check(lib().ptc_set_timing(self._h, 1 if on else 0), self)

    def stream(self) -> int:
        return int(lib().ptc_stream(self._h) or 0)

    def close(self) -> None:
        if self._h:
            # tables and batch plans before tensors (they reference tensors)
            kids = sorted(list(self._children), key=lambda o: 0 if isinstance(o, (HashTable, BatchPlan)) else 1)
            for obj in kids:
                try:
                    obj.free()
                except Exception:
                    pass
            lib().ptc_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context()
    return _default_ctx


class DeviceTensor:
    """A COO tensor resident in HBM (ptc_dtensor)."""

    def __init__(self, ctx: Context, handle: C.c_void_p, keep=None):
        self.ctx = ctx
        self._h = handle
        self._keep = keep
        ctx.adopt(self)

    @classmethod
    def upload(cls, t: Tensor, ctx: Optional[Context] = None) -> "DeviceTensor":
        ctx = ctx or default_context()
        h = C.c_void_p()
        v = t.view()
        _check(lib().ptc_dtensor_upload(ctx.handle, C.byref(v), C.byref(h)), ctx)
        return cls(ctx, h)

    @classmethod
    def upload_f32(cls, t: Tensor, ctx: Optional[Context] = None) -> "DeviceTensor":
        """Upload with the values rounded to FP32 (nearest even)."""
        ctx = ctx or default_context()
        h = C.c_void_p()
        v = t.view()
        v32 = np.ascontiguousarray(t.val_array, dtype=np.float32)
        _check(lib().ptc_dtensor_upload_f32(ctx.handle, C.byref(v), v32.ctypes.data_as(C.POINTER(C.c_float)),
                                            C.byref(h)), ctx)
        return cls(ctx, h, keep=v32)

    def info(self):
        d, n, nk = C.c_uint32(), C.c_uint64(), C.c_uint32()
        _check(lib().ptc_dtensor_info(self._h, C.byref(d), C.byref(n), C.byref(nk)))
        return int(d.value), int(n.value), int(nk.value)

    def set_keys(self, num_key_dimensions: int, dim_tc) -> None:
        perm = np.ascontiguousarray(dim_tc, dtype=np.uint32)
        _check(lib().ptc_dtensor_set_keys(self._h, num_key_dimensions, _p32(perm)), self.ctx)

    def download(self) -> Tensor:
        d, n, nk = self.info()
        h = np.empty((n, d), dtype=np.uint32)
        v = np.empty(n, dtype=np.float64)
        dims = np.zeros(MAX_MODES, dtype=np.uint32)
        _check(lib().ptc_dtensor_download(self.ctx.handle, self._h, _p32(h), _pd(v), _p32(dims)), self.ctx)
        return Tensor(h, v, dims[:d].copy(), np.arange(d, dtype=np.uint32), nk)

    def set_values(self, vals) -> None:
        """Overwrite the FP64 values in place (same coordinates)."""
        v = np.ascontiguousarray(vals, dtype=np.float64)
        _check(lib().ptc_dtensor_set_values(self.ctx.handle, self._h, _pd(v)), self.ctx)

    def download_f32(self, with_lo: bool = False):
        """FP32 values; with_lo=True also returns the lo part of an FP32
        contraction result (hi + lo ~ the sum to twice FP32 precision)."""
        d, n, nk = self.info()
        v = np.empty(n, dtype=np.float32)
        lo = np.empty(n, dtype=np.float32) if with_lo else None
        pf = C.POINTER(C.c_float)
        _check(lib().ptc_dtensor_download_f32(self.ctx.handle, self._h, v.ctypes.data_as(pf),
                                              lo.ctypes.data_as(pf) if with_lo else None), self.ctx)
        return (v, lo) if with_lo else v

    def free(self) -> None:
        if self._h:
            lib().ptc_dtensor_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HashTable:
    """Device key table over a tensor's key modes (buildHashTable)."""

    def __init__(self, ctx: Context, handle: C.c_void_p, tensor: DeviceTensor):
        self.ctx = ctx
        self._h = handle
        self.tensor = tensor
        ctx.adopt(self)

    def info(self):
        k, n = C.c_uint64(), C.c_uint64()
        _check(lib().ptc_hashtable_info(self._h, C.byref(k), C.byref(n)))
        return int(k.value), int(n.value)

    def find(self, keys: np.ndarray, with_ids: bool = True):
        """Probe (n, nk) keys: slice lengths and, optionally, the original
        nonzero ids of each slice (find_element_new, fksCuckoo.hpp:867)."""
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        if keys.ndim == 1:
            keys = keys.reshape(1, -1)
        n = keys.shape[0]
        lens = np.zeros(n, dtype=np.uint64)
        offs = np.zeros(n + 1, dtype=np.uint64)
        _check(lib().ptc_hashtable_probe(self.ctx.handle, self._h, n, _p32(keys), _p64(lens), _p64(offs),
                                         None, 0), self.ctx)
        if not with_ids:
            return lens
        total = int(offs[-1])
        ids = np.zeros(max(total, 1), dtype=np.uint32)
        _check(lib().ptc_hashtable_probe(self.ctx.handle, self._h, n, _p32(keys), _p64(lens), _p64(offs),
                                         _p32(ids), total), self.ctx)
        return [ids[int(offs[i]):int(offs[i + 1])] for i in range(n)]

    def free(self) -> None:
        if self._h:
            lib().ptc_hashtable_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _as_device(t, ctx: Context) -> DeviceTensor:
    return t if isinstance(t, DeviceTensor) else DeviceTensor.upload(t, ctx)


def buildHashTable(T, ctx: Optional[Context] = None, grouping: int = GROUP_AUTO) -> HashTable:
    """Group T's nonzeros by its key modes (T.dim_tc[:num_key_dimensions])
    on the device (fksCuckoo.hpp:1044-1060 with reorder_tensor); grouping
    GROUP_HASH, GROUP_SORT or GROUP_PARTITION picks the method
    (ptc_hashtable_build_ex)."""
    ctx = ctx or default_context()
    dT = _as_device(T, ctx)
    h = C.c_void_p()
    _check(lib().ptc_hashtable_build_ex(ctx.handle, dT._h, grouping, C.byref(h)), ctx)
    return HashTable(ctx, h, dT)


def options(method: int = METHOD_TWO_PASS, value_type: int = VAL_F64, acc_mode: int = ACC_AUTO,
            dense_max_cols: int = 0, lds_hash_max: int = 0, chunk_flops: int = 0, predict_scale: float = 0.0,
            mlp=None, deterministic: int = 0, block_threads: int = 0) -> Options:
    o = Options()
    lib().ptc_options_default(C.byref(o))
    o.method = method
    o.value_type = value_type
    o.acc_mode = acc_mode
    o.dense_max_cols = dense_max_cols
    o.lds_hash_max = lds_hash_max
    o.chunk_flops = chunk_flops
    if predict_scale:
        o.predict_scale = predict_scale
    o.mlp = mlp._h if mlp is not None else None
    o.deterministic = deterministic
    o.block_threads = block_threads
    return o


def tc_fksCuckoo(HT_A: HashTable, HT_B: HashTable, A=None, B=None, *, download: bool = True,
                 stats: Optional[Stats] = None, **opts):
    """O = A x B (fksCuckooTC.hpp:80).  A free-first, B contracted-first; the
    tables must have been built on A and B."""
    ctx = HT_A.ctx
    o = options(**opts)
    out = C.c_void_p()
    st = stats if stats is not None else Stats()
    _check(lib().ptc_tc_fksCuckoo(ctx.handle, HT_A._h, HT_B._h, HT_A.tensor._h, HT_B.tensor._h, C.byref(o),
                                  C.byref(out), C.byref(st)), ctx)
    dO = DeviceTensor(ctx, out)
    return dO.download() if download else dO


def tc_fksCuckoo_keyed(HT_A: HashTable, HT_B: HashTable, num_key_dimensions: int, dim_tc, *,
                       stats: Optional[Stats] = None, **opts):
    """One step of a contraction chain (ptc_tc_fksCuckoo_keyed): O = A x B
    as tc_fksCuckoo computes it, with O's key split set to
    (num_key_dimensions, dim_tc) and O's key table for the next contraction.
    Returns (DeviceTensor O, HashTable of O, from_rows): from_rows is True
    when the table came from the output's row grouping (key modes = O's row
    modes) rather than a build."""
    ctx = HT_A.ctx
    o = options(**opts)
    perm = np.ascontiguousarray(dim_tc, dtype=np.uint32)
    out, tab, fr = C.c_void_p(), C.c_void_p(), C.c_int(0)
    st = stats if stats is not None else Stats()
    _check(lib().ptc_tc_fksCuckoo_keyed(ctx.handle, HT_A._h, HT_B._h, HT_A.tensor._h, HT_B.tensor._h, C.byref(o),
                                        num_key_dimensions, _p32(perm), C.byref(out), C.byref(tab), C.byref(fr),
                                        C.byref(st)), ctx)
    dO = DeviceTensor(ctx, out)
    return dO, HashTable(ctx, tab, dO), bool(fr.value)


def contract(A, B, ctx: Optional[Context] = None, *, download: bool = True, stats: Optional[Stats] = None,
             **opts):
    """Both tensors with .meta-style keys (contracted modes first); the
    driver path of main.cpp:59-118 (A_copy, HT_B then HT_A, contract)."""
    ctx = ctx or default_context()
    dA, dB = _as_device(A, ctx), _as_device(B, ctx)
    o = options(**opts)
    out = C.c_void_p()
    st = stats if stats is not None else Stats()
    _check(lib().ptc_contract(ctx.handle, dA._h, dB._h, C.byref(o), C.byref(out), C.byref(st)), ctx)
    dO = DeviceTensor(ctx, out)
    return dO.download() if download else dO


def _handles(ts, ctx):
    devs = [_as_device(t, ctx) for t in ts]
    arr = (C.c_void_p * len(devs))(*[d._h.value for d in devs])
    return devs, arr


class BatchPlan:
    """A batch of contractions planned on the device (ptc_batch_plan): run()
    launches the captured numeric graph; after changing input values in
    place (DeviceTensor.set_values) run() recomputes every output."""

    def __init__(self, As, Bs, ctx: Optional[Context] = None, **opts):
        if len(As) != len(Bs) or not As:
            raise ValueError("need equally many A and B tensors")
        self.ctx = ctx or default_context()
        self._A, a = _handles(As, self.ctx)
        self._B, b = _handles(Bs, self.ctx)
        o = options(**opts)
        self.stats = BatchStats()
        h = C.c_void_p()
        _check(lib().ptc_batch_plan(self.ctx.handle, len(As), a, b, C.byref(o), C.byref(h), C.byref(self.stats)),
               self.ctx)
        self._h = h
        self.n = len(As)
        self.ctx.adopt(self)

    @property
    def inputs(self):
        return self._A, self._B

    def run(self) -> BatchStats:
        _check(lib().ptc_batch_run(self._h, C.byref(self.stats)), self.ctx)
        return self.stats

    def output(self, j: int) -> Tensor:
        h = C.c_void_p()
        _check(lib().ptc_batch_output(self._h, j, C.byref(h)), self.ctx)
        t = DeviceTensor(self.ctx, h)
        try:
            return t.download()
        finally:
            t._h = C.c_void_p()  # owned by the plan

    def free(self) -> None:
        if self._h:
            lib().ptc_batch_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def autotune_batch(As, Bs, ctx: Optional[Context] = None, **base_opts):
    """ptc_batch_autotune: returns (tuned options as keyword arguments, the
    tuner's report)."""
    ctx = ctx or default_context()
    keep_a, a = _handles(As, ctx)
    keep_b, b = _handles(Bs, ctx)
    base = options(**base_opts)
    best = Options()
    rep = TuneReport()
    _check(lib().ptc_batch_autotune(ctx.handle, len(As), a, b, C.byref(base), C.byref(best), C.byref(rep)), ctx)
    del keep_a, keep_b
    tuned = dict(base_opts, dense_max_cols=best.dense_max_cols, lds_hash_max=best.lds_hash_max,
                 chunk_flops=best.chunk_flops)
    return tuned, rep.as_dict()


def contract_batched(As, Bs, ctx: Optional[Context] = None, *, download: bool = True,
                     stats: Optional[BatchStats] = None, **opts):
    """Every (A[j], B[j]) contracted as contract() would, as one batch
    (ptc_contract_batched)."""
    ctx = ctx or default_context()
    keep_a, a = _handles(As, ctx)  # the device tensors must outlive the call
    keep_b, b = _handles(Bs, ctx)
    o = options(**opts)
    outs = (C.c_void_p * len(As))()
    st = stats if stats is not None else BatchStats()
    _check(lib().ptc_contract_batched(ctx.handle, len(As), a, b, C.byref(o), outs, C.byref(st)), ctx)
    del keep_a, keep_b
    res = [DeviceTensor(ctx, C.c_void_p(x)) for x in outs]
    return [r.download() for r in res] if download else res


def matricize_spGEMM(A, B, ctx: Optional[Context] = None, *, download: bool = True,
                     stats: Optional[Stats] = None):
    """matricizedspGEMM.hpp:172 -- both tensors contracted-first.  The
    column-wise product of cs_multiply (cs_multiply.c:6-38): the output comes
    column by column (B's free tuples), each column's entries over the rows."""
    ctx = ctx or default_context()
    dA, dB = _as_device(A, ctx), _as_device(B, ctx)
    out = C.c_void_p()
    st = stats if stats is not None else Stats()
    _check(lib().ptc_matricize_spGEMM(ctx.handle, dA._h, dB._h, C.byref(out), C.byref(st)), ctx)
    dO = DeviceTensor(ctx, out)
    return dO.download() if download else dO


def matricize_sort_spGEMM(A, B, ctx: Optional[Context] = None) -> Tensor:
    """matricizedspGEMM_sorting.hpp:194 -- same product."""
    ctx = ctx or default_context()
    dA, dB = _as_device(A, ctx), _as_device(B, ctx)
    out = C.c_void_p()
    _check(lib().ptc_matricize_sort_spGEMM(ctx.handle, dA._h, dB._h, C.byref(out), None), ctx)
    return DeviceTensor(ctx, out).download()


# --------------------------------------------------------- synthetic inputs


def gen_coo(seed: int, dims, nnz: int, dist: int = DIST_ZIPF, alpha: float = 1.1):
    """Seeded unique-coordinate COO (SplitMix64/xoshiro256**, truncated Zipf
    by inverse CDF or uniform; values U(0,1])."""
    dims_a = np.ascontiguousarray(dims, dtype=np.uint32)
    d = dims_a.shape[0]
    h = np.empty((nnz, d), dtype=np.uint32)
    v = np.empty(nnz, dtype=np.float64)
    rc = lib().ptc_gen_coo(seed, d, _p32(dims_a), nnz, dist, alpha, _p32(h), _pd(v))
    if rc != PTC_OK:
        raise PtcError(rc, "ptc_gen_coo")
    return h, v


def gen_job_shape(seed: int, dim_lo=64, dim_hi=20000, nnz_lo=1000, nnz_hi=1000000):
    d = C.c_uint32()
    dims = np.zeros(MAX_MODES, dtype=np.uint32)
    nnz = C.c_uint64()
    rc = lib().ptc_gen_job_shape(seed, dim_lo, dim_hi, nnz_lo, nnz_hi, C.byref(d), _p32(dims), C.byref(nnz))
    if rc != PTC_OK:
        raise PtcError(rc, "ptc_gen_job_shape")
    return int(d.value), dims[: d.value].copy(), int(nnz.value)


def self_contraction_inputs(h: np.ndarray, v: np.ndarray, contracted):
    """(A_copy, B) of main.cpp's self-contraction: the .meta permutation has
    the contracted modes first (4d.meta), A_copy rotates it to free-first,
    B is A (sparse_tensor_generator.hpp:4-88)."""
    d = h.shape[1]
    contracted = [int(m) for m in contracted]
    free = [m for m in range(d) if m not in contracted]
    perm = np.array(contracted + free, dtype=np.uint32)
    A = Tensor(h, v, None, perm, len(contracted))
    return make_free_first(A), get_sparse_tensor(A)


def learned_available() -> bool:
    """True when a trained row-size model is loaded (or loadable) as the
    LEARNED method's default."""
    L = lib()
    L.ptc_learned_default.restype = C.c_void_p
    if L.ptc_learned_default():
        return True
    return load_default_model()


# ------------------------------------------------------- learned row sizes

MLP_IN = 6
MLP_HIDDEN = 64
MLP_NPARAM = MLP_HIDDEN * MLP_IN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN + 1
DEFAULT_MODEL = _PKG / "models" / "rowsize_mlp.bin"


def _mlp_sigs(L):
    vp, u32, u64, i32, f32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_float
    PF = C.POINTER(C.c_float)
    sig = {
        "ptc_mlp_create": (i32, [vp, u64, C.POINTER(vp)]),
        "ptc_mlp_free": (None, [vp]),
        "ptc_mlp_num_params": (i32, []),
        "ptc_mlp_get_params": (i32, [vp, PF]),
        "ptc_mlp_set_params": (i32, [vp, PF]),
        "ptc_mlp_grad": (i32, [vp, vp, PF, PF, PF, u64, PF, PF]),
        "ptc_mlp_train": (i32, [vp, vp, PF, PF, PF, u64, u32, u32, f32, u64, PF]),
        "ptc_mlp_predict": (i32, [vp, vp, PF, PF, u64, PF]),
        "ptc_mlp_save": (i32, [vp, C.c_char_p]),
        "ptc_mlp_load": (i32, [vp, C.c_char_p, C.POINTER(vp)]),
        "ptc_learned_set_default": (None, [vp]),
        "ptc_learned_default": (vp, []),
        "ptc_row_profile": (i32, [vp, vp, vp, vp, vp, PF, PF, C.POINTER(C.c_uint32), u64, C.POINTER(C.c_uint64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _pf(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class RowSizeModel:
    """The row-size network (ptc_mlp): 6 -> 64 -> 64 -> 1, trained on the GPU."""

    def __init__(self, ctx: Optional[Context] = None, seed: int = 1, handle=None):
        self.ctx = ctx or default_context()
        _mlp_sigs(lib())
        if handle is not None:
            self._h = handle
        else:
            self._h = C.c_void_p()
            _check(lib().ptc_mlp_create(self.ctx.handle, seed, C.byref(self._h)), self.ctx)
        self.ctx.adopt(self)

    @classmethod
    def load(cls, path, ctx: Optional[Context] = None) -> "RowSizeModel":
        ctx = ctx or default_context()
        _mlp_sigs(lib())
        h = C.c_void_p()
        _check(lib().ptc_mlp_load(ctx.handle, str(path).encode(), C.byref(h)), ctx)
        return cls(ctx, handle=h)

    def save(self, path) -> None:
        Path(path).parent.mkdir(parents=True, exist_ok=True)
        _check(lib().ptc_mlp_save(self._h, str(path).encode()), self.ctx)

    def params(self) -> np.ndarray:
        w = np.zeros(MLP_NPARAM, dtype=np.float32)
        _check(lib().ptc_mlp_get_params(self._h, _pf(w)), self.ctx)
        return w

    def set_params(self, w) -> None:
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.shape == (MLP_NPARAM,)
        _check(lib().ptc_mlp_set_params(self._h, _pf(w)), self.ctx)

    @staticmethod
    def _arrs(X, ub, y=None):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, MLP_IN)
        ub = np.ascontiguousarray(ub, dtype=np.float32).reshape(-1)
        y = None if y is None else np.ascontiguousarray(y, dtype=np.float32).reshape(-1)
        return X, ub, y

    def grad(self, X, ub, y):
        X, ub, y = self._arrs(X, ub, y)
        g = np.zeros(MLP_NPARAM, dtype=np.float32)
        loss = C.c_float()
        _check(lib().ptc_mlp_grad(self.ctx.handle, self._h, _pf(X), _pf(ub), _pf(y), X.shape[0], _pf(g),
                                  C.byref(loss)), self.ctx)
        return g, float(loss.value)

    def train(self, X, ub, y, epochs=20, batch=4096, lr=1e-3, seed=7):
        X, ub, y = self._arrs(X, ub, y)
        losses = np.zeros(max(epochs, 1), dtype=np.float32)
        _check(lib().ptc_mlp_train(self.ctx.handle, self._h, _pf(X), _pf(ub), _pf(y), X.shape[0], epochs, batch, lr,
                                   seed, _pf(losses)), self.ctx)
        return losses[:epochs]

    def predict(self, X, ub):
        X, ub, _ = self._arrs(X, ub)
        p = np.zeros(X.shape[0], dtype=np.float32)
        _check(lib().ptc_mlp_predict(self.ctx.handle, self._h, _pf(X), _pf(ub), X.shape[0], _pf(p)), self.ctx)
        return p

    def set_default(self) -> None:
        lib().ptc_learned_set_default(self._h)

    def free(self) -> None:
        if self._h:
            lib().ptc_mlp_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def row_profile(HT_A: HashTable, HT_B: HashTable):
    """(X, ub, nnz) per output row: learned features and the exact symbolic
    counts (training data of the row-size network)."""
    _mlp_sigs(lib())
    ctx = HT_A.ctx
    nrows = C.c_uint64()
    k, _ = HT_A.info()
    X = np.zeros((max(k, 1), MLP_IN), dtype=np.float32)
    ub = np.zeros(max(k, 1), dtype=np.float32)
    nnz = np.zeros(max(k, 1), dtype=np.uint32)
    _check(lib().ptc_row_profile(ctx.handle, HT_A._h, HT_B._h, HT_A.tensor._h, HT_B.tensor._h, _pf(X), _pf(ub),
                                 nnz.ctypes.data_as(C.POINTER(C.c_uint32)), X.shape[0], C.byref(nrows)), ctx)
    n = int(nrows.value)
    return X[:n], ub[:n], nnz[:n]


_default_model = None


def load_default_model(ctx: Optional[Context] = None) -> bool:
    """Load partc_amd/models/rowsize_mlp.bin (if present) as the model of the
    LEARNED method."""
    global _default_model
    if _default_model is not None:
        return True
    if not DEFAULT_MODEL.exists():
        return False
    _default_model = RowSizeModel.load(DEFAULT_MODEL, ctx)
    _default_model.set_default()
    return True
