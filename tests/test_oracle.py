"""Pins the oracle (the CPU restatement of tc_fksCuckoo and matricize_spGEMM)
before anything is compared against it.

The reference ships no tests or fixtures for this path, so the oracle is
pinned by:
  * golden fixtures written by tests/golden/make_golden.py from two
    formulations independent of the oracle (dense matrix product; a plain
    loop in the reference's summation order, which pins values bit-exactly);
  * the reference's own CXSparse product (oracle/_ref/libcxsparse.so, built
    from the reference's sources) and scipy's sparse product on matricized
    tensors;
  * closed forms and invariants: F = sum_k |A_k| |B_k|; a self-contraction's
    coordinate set is symmetric; a single-nonzero row is its scaled slice;
  * fastmod against % on the divisor ranges of its upstream unit test
    (extern/fastmod-master/tests/unit.c:217-235);
  * planted mistakes in the oracle that these checks must catch.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import partc_amd as P
from oracle import orc
from tests.helpers import golden_cases, max_rel, meta_tensors, sort_coo, tc_inputs

CASES = golden_cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_golden(name):
    case = CASES[name]
    A, B = tc_inputs(case)
    h, v, info = orc.tc_fksCuckoo(A, B)
    h, v = sort_coo(h, v)
    assert np.array_equal(h, case["coords"])
    # bit-exact against the reference's summation order
    assert np.array_equal(v, case["vals_ordered"])
    assert max_rel(v, case["vals_dense"]) <= 1e-12
    assert info["flops"] == int(case["flops"][0])


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_driver_path_matches_golden(name):
    """orc_contract_meta (main.cpp:68-118: A_copy made free-first from the
    .meta tensor) gives the golden product, in the same emission order as
    tc_fksCuckoo on the caller-made A_copy."""
    case = CASES[name]
    A, B = meta_tensors(case)
    h, v, info = orc.contract_meta(A, B)
    hs, vs = sort_coo(h, v)
    assert np.array_equal(hs, case["coords"])
    assert np.array_equal(vs, case["vals_ordered"])
    assert info["flops"] == int(case["flops"][0])
    h2, v2, _ = orc.tc_fksCuckoo(*tc_inputs(case))
    assert np.array_equal(h, h2) and np.array_equal(v, v2)


def test_oracle_threads_agree_with_serial():
    """Oracle contractions on several host threads at once (the RNG stream of
    the hash keys is per thread) give what they give one at a time."""
    from concurrent.futures import ThreadPoolExecutor

    ins = []
    for i in range(6):
        h, v = P.gen_coo(70 + i, [80, 60, 50], 6000, P.DIST_ZIPF)
        ins.append(P.self_contraction_inputs(h, v, [1, 2]))
    serial = [sort_coo(*orc.tc_fksCuckoo(A, B)[:2]) for A, B in ins]
    with ThreadPoolExecutor(6) as ex:
        par = list(ex.map(lambda ab: sort_coo(*orc.tc_fksCuckoo(*ab)[:2]), ins))
    for (h1, v1), (h2, v2) in zip(serial, par):
        assert np.array_equal(h1, h2)
        assert np.array_equal(v1, v2)


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matricized_matches_golden(name):
    case = CASES[name]
    A, B = meta_tensors(case)
    h, v, _ = orc.matricize_spGEMM(A, B)
    h, v = sort_coo(h, v)
    assert np.array_equal(h, case["coords"])
    assert max_rel(v, case["vals_dense"]) <= 1e-12


def _matricize(A, B):
    """CSC pieces of A_mat (rows = A free tuples) and B_mat (cols = B free
    tuples) with a shared contracted-tuple numbering; plus decoders."""
    c = int(A.num_key_dimensions)
    ka = [tuple(r) for r in A.hedges_array[:, A.dim_tc[:c]]]
    kb = [tuple(r) for r in B.hedges_array[:, B.dim_tc[:c]]]
    kid = {}
    for t in ka + kb:
        kid.setdefault(t, len(kid))
    ra = [tuple(r) for r in A.hedges_array[:, A.dim_tc[c:]]]
    cb = [tuple(r) for r in B.hedges_array[:, B.dim_tc[c:]]]
    rid, cid = {}, {}
    for t in ra:
        rid.setdefault(t, len(rid))
    for t in cb:
        cid.setdefault(t, len(cid))
    Am = sp.csc_matrix((A.val_array, ([rid[t] for t in ra], [kid[t] for t in ka])), shape=(len(rid), len(kid)))
    Bm = sp.csc_matrix((B.val_array, ([kid[t] for t in kb], [cid[t] for t in cb])), shape=(len(kid), len(cid)))
    rinv = np.array(list(rid.keys()), dtype=np.uint32).reshape(len(rid), -1)
    cinv = np.array(list(cid.keys()), dtype=np.uint32).reshape(len(cid), -1)
    return Am, Bm, rinv, cinv


def _coo_from(rows, cols, vals, rinv, cinv):
    h = np.concatenate([rinv[rows], cinv[cols]], axis=1).astype(np.uint32)
    return sort_coo(h, np.asarray(vals, dtype=np.float64))


@pytest.mark.skipif(not orc.cxsparse_available(), reason="reference CXSparse not built (no reference tree)")
@pytest.mark.parametrize("name", sorted(CASES) + ["C1"])
def test_oracle_matches_reference_cxsparse(name):
    if name == "C1":
        h, v = P.gen_coo(1, [200, 300, 400], 50000, P.DIST_UNIFORM)
        A = P.Tensor(h, v, None, np.array([1, 2, 0], np.uint32), 2)
        B = P.get_sparse_tensor(A)
    else:
        A, B = meta_tensors(CASES[name])
    Am, Bm, rinv, cinv = _matricize(A, B)
    Am.sort_indices()
    Bm.sort_indices()
    p, i, x = orc.cxsparse_multiply(Am.indptr, Am.indices, Am.data, Am.shape[0], Am.shape[1], Bm.indptr,
                                    Bm.indices, Bm.data, Bm.shape[1])
    cols = np.repeat(np.arange(Bm.shape[1]), np.diff(p))
    eh, ev = _coo_from(i, cols, x, rinv, cinv)
    oh, ov, _ = orc.tc_fksCuckoo(P.make_free_first(A), B)
    oh, ov = sort_coo(oh, ov)
    assert np.array_equal(oh, eh)
    assert max_rel(ov, ev) <= 1e-12


@pytest.mark.parametrize("seed,zipf", [(11, False), (12, True), (13, True)])
def test_oracle_matches_scipy(seed, zipf):
    h, v = P.gen_coo(seed, [60, 70, 80, 9], 20000, P.DIST_ZIPF if zipf else P.DIST_UNIFORM)
    A = P.Tensor(h, v, None, np.array([1, 3, 0, 2], np.uint32), 2)
    B = P.get_sparse_tensor(A)
    Am, Bm, rinv, cinv = _matricize(A, B)
    Cm = (Am @ Bm).tocoo()
    eh, ev = _coo_from(Cm.row, Cm.col, Cm.data, rinv, cinv)
    oh, ov, info = orc.tc_fksCuckoo(P.make_free_first(A), B)
    oh, ov = sort_coo(oh, ov)
    assert np.array_equal(oh, eh)
    assert max_rel(ov, ev) <= 1e-12
    # closed form F = sum_k |A_k||B_k|
    assert info["flops"] == orc.flops(A, B)
    keys = h[:, [1, 3]].astype(np.int64)
    _, cnt = np.unique(keys[:, 0] * 100 + keys[:, 1], return_counts=True)
    assert info["flops"] == int((cnt.astype(np.int64) ** 2).sum())


def test_oracle_self_contraction_invariants():
    h, v = P.gen_coo(5, [120, 50, 60], 30000, P.DIST_ZIPF)
    A, B = P.self_contraction_inputs(h, v, [1, 2])
    oh, ov, info = orc.tc_fksCuckoo(A, B)
    # unique coordinates
    assert np.unique(oh, axis=0).shape[0] == oh.shape[0]
    # symmetric coordinate set, values symmetric within rounding
    sh, sv = sort_coo(oh, ov)
    th, tv = sort_coo(oh[:, ::-1].copy(), ov)
    assert np.array_equal(sh, th)
    assert max_rel(sv, tv) <= 1e-12
    assert np.all(ov > 0)
    # nnz <= min(F, rows * cols)
    assert oh.shape[0] <= min(info["flops"], 120 * 120)


def test_oracle_single_nonzero_row_is_scaled_slice():
    # row 0 of A has exactly one nonzero: its output row is A[0,k] * B[k,:]
    h = np.array([[0, 1, 2], [1, 1, 2], [1, 3, 3], [2, 1, 2], [2, 3, 3], [3, 0, 0]], np.uint32)
    v = np.array([0.5, 0.25, 0.75, 0.125, 0.375, 0.9])
    A, B = P.self_contraction_inputs(h, v, [1, 2])
    oh, ov, _ = orc.tc_fksCuckoo(A, B)
    sel = oh[:, 0] == 0
    got = dict(zip(oh[sel, 1].tolist(), ov[sel].tolist()))
    assert got == {0: 0.5 * 0.5, 1: 0.5 * 0.25, 2: 0.5 * 0.125}


def test_oracle_empty_contraction_reports_error():
    A = P.Tensor(np.array([[0, 1], [1, 2]], np.uint32), np.array([1.0, 2.0]), None, np.array([0, 1], np.uint32), 1)
    B = P.Tensor(np.array([[5, 1], [6, 2]], np.uint32), np.array([1.0, 2.0]), None, np.array([0, 1], np.uint32), 1)
    with pytest.raises(orc.OracleError) as e:
        orc.tc_fksCuckoo(P.make_free_first(A), B)
    assert e.value.code == -4


@pytest.mark.parametrize("side", ["A", "B"])
def test_oracle_empty_operand_reports_error(side):
    """An operand with no nonzero has no contributing pair either."""
    h, _ = P.gen_coo(3, [300, 50], 200, P.DIST_UNIFORM)
    A = P.Tensor(h, np.ones(h.shape[0]), np.array([300, 50], np.uint32), np.array([0, 1], np.uint32), 1)
    h, _ = P.gen_coo(4, [50, 400], 200, P.DIST_UNIFORM)
    B = P.Tensor(h, np.ones(h.shape[0]), np.array([50, 400], np.uint32), np.array([0, 1], np.uint32), 1)
    empty = lambda T: P.Tensor(np.zeros((0, 2), np.uint32), np.zeros(0), T.dimensions_array, T.dim_tc, 1)
    if side == "A":
        A = empty(A)
    else:
        B = empty(B)
    with pytest.raises(orc.OracleError) as e:
        orc.tc_fksCuckoo(A, B)
    assert e.value.code == -4


def test_fastmod_matches_modulo():
    rng = np.random.default_rng(3)
    nums = np.concatenate([np.arange(0, 4096), rng.integers(0, 2 ** 32, 4096),
                           np.array([2 ** 32 - 1, 2 ** 32 - 2, 2 ** 31, 2 ** 31 - 1])]).astype(np.uint64)
    for d in list(range(1, 9)) + list(range(0xFFFFFFF8, 0x100000000)) + [3, 7, 2481, 3101609, 2147483647]:
        for a in nums[::7]:
            assert orc.fastmod_u32(int(a), d) == int(a) % d


@pytest.mark.parametrize("plant", [1, 2, 3, 4])
def test_planted_oracle_mistakes_are_caught(plant):
    """Each planted mistake must make the golden comparison fail."""
    caught = False
    try:
        orc.lib().orc_set_plant(plant)
        for name, case in CASES.items():
            A, B = tc_inputs(case)
            try:
                h, v, _ = orc.tc_fksCuckoo(A, B)
            except orc.OracleError:
                caught = True
                break
            h, v = sort_coo(h, v)
            if h.shape != case["coords"].shape or not np.array_equal(h, case["coords"]) or not np.array_equal(
                    v, case["vals_ordered"]):
                caught = True
                break
    finally:
        orc.lib().orc_set_plant(0)
    assert caught, f"planted mistake {plant} went unnoticed"
