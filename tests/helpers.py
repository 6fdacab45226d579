"""Shared helpers of the test suite: golden fixtures and input builders."""
from __future__ import annotations

from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden" / "contractions.npz"


def golden_cases():
    z = np.load(GOLDEN, allow_pickle=False)
    names = sorted({k.split("/")[0] for k in z.files})
    out = {}
    for n in names:
        out[n] = {k.split("/")[1]: z[k] for k in z.files if k.startswith(n + "/")}
    return out


def meta_tensors(case):
    """(A, B) as the reference's driver holds them after populate_meta_info:
    .meta permutation (contracted first), num_key_dimensions = c."""
    import partc_amd as P

    c = int(case["c"][0])
    A = P.Tensor(case["Ah"], case["Av"], case["Adims"], case["Aperm"], c)
    B = P.Tensor(case["Bh"], case["Bv"], case["Bdims"], case["Bperm"], c)
    return A, B


def tc_inputs(case):
    """(A_copy, B) in tc_fksCuckoo's layout (A free-first, B contracted-first)."""
    import partc_amd as P

    A, B = meta_tensors(case)
    return P.make_free_first(A), B


def sort_coo(h, v):
    if h.shape[0] == 0:
        return h, v
    o = np.lexsort(h.T[::-1])
    return h[o], v[o]


def max_rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return 0.0
    den = np.maximum(np.abs(b), np.finfo(np.float64).tiny)
    return float(np.max(np.abs(a - b) / den))


def random_case(rng, d, dims, nnz, contracted, zipf=False, seed=None):
    """A random self-contraction in tc layout using the engine's generator."""
    import partc_amd as P

    h, v = P.gen_coo(int(seed if seed is not None else rng.integers(1 << 40)), dims, nnz,
                     P.DIST_ZIPF if zipf else P.DIST_UNIFORM)
    return P.self_contraction_inputs(h, v, contracted)
