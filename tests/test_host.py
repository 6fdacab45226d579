"""Host-side checks that need no GPU: the C ABI library loads and exports
everything include/partc.h declares, the .meta parser, the generator, and the
Tensor mirror of the reference's driver conventions."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import partc_amd as P

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = []
    for hdr in sorted((ROOT / "include").glob("*.h")):
        text = hdr.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(ptc_\w+)\s*\(", text, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("ptc_tc_fksCuckoo", "ptc_hashtable_build", "ptc_meta_parse", "ptc_contract",
                 "ptc_matricize_spGEMM", "ptc_matricize_sort_spGEMM", "ptc_dtensor_upload"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(P.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_meta_parse(tmp_path):
    f = tmp_path / "4d.meta"
    f.write_text("#line1: number of contraction modes, c\n#line2: permutation\n2\n1 2 0 3\n")
    t = P.Tensor(np.zeros((1, 4), np.uint32), np.ones(1))
    t.populate_meta_info(str(f))
    assert t.num_key_dimensions == 2
    assert t.dim_tc.tolist() == [1, 2, 0, 3]
    bad = tmp_path / "bad.meta"
    bad.write_text("a\nb\n2\n1 1 0 3\n")
    with pytest.raises(P.PtcError):
        t.populate_meta_info(str(bad))
    with pytest.raises(P.PtcError):
        t.populate_meta_info(str(tmp_path / "missing.meta"))


def test_free_first_matches_driver_convention():
    # main.cpp:74-84: A_copy.nk = d - c; dim_tc = [pi[c..d), pi[0..c)]
    A = P.Tensor(np.zeros((1, 4), np.uint32), np.ones(1), None, np.array([1, 2, 0, 3], np.uint32), 2)
    Ac = P.make_free_first(A)
    assert Ac.num_key_dimensions == 2
    assert Ac.dim_tc.tolist() == [0, 3, 1, 2]
    B = P.get_sparse_tensor(A)
    assert B.num_key_dimensions == 2 and B.dim_tc.tolist() == [1, 2, 0, 3]
    assert np.array_equal(B.hedges_array, A.hedges_array)


def test_generator_is_deterministic_and_unique():
    h1, v1 = P.gen_coo(2, [2482, 2862, 14036, 17], 20000, P.DIST_ZIPF)
    h2, v2 = P.gen_coo(2, [2482, 2862, 14036, 17], 20000, P.DIST_ZIPF)
    assert np.array_equal(h1, h2) and np.array_equal(v1, v2)
    assert np.unique(h1, axis=0).shape[0] == h1.shape[0]
    assert np.all(v1 > 0) and np.all(v1 <= 1)
    assert np.all(h1.max(axis=0) < np.array([2482, 2862, 14036, 17]))
    # Zipf: index 0 is the most frequent in every mode
    for j in range(4):
        counts = np.bincount(h1[:, j])
        assert counts.argmax() == 0
    # pinned stream (guards the generator across machines)
    assert h1[:3].tolist() == PINNED_HEAD
    assert float(v1[:1000].sum()).hex() == PINNED_VSUM


PINNED_HEAD = [[0, 122, 1, 6], [1, 58, 2, 3], [138, 5, 164, 14]]
PINNED_VSUM = "0x1.f804e12f02497p+8"


def test_generator_uniform_and_errors():
    h, v = P.gen_coo(1, [200, 300, 400], 1000, P.DIST_UNIFORM)
    assert h.shape == (1000, 3)
    with pytest.raises(P.PtcError):
        P.gen_coo(1, [2, 2], 5, P.DIST_UNIFORM)  # more nnz than coordinates


def test_job_shapes():
    seen3 = seen4 = 0
    for i in range(64):
        d, dims, nnz = P.gen_job_shape(1000 + i)
        assert d in (3, 4)
        assert np.all(dims >= 64) and np.all(dims <= 20000)
        assert 1 <= nnz <= 1000000
        seen3 += d == 3
        seen4 += d == 4
    assert seen3 and seen4


def test_no_device_is_an_error_not_a_fallback():
    import os

    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    with pytest.raises(P.PtcError) as e:
        P.Context(0)
    assert e.value.code == -8
