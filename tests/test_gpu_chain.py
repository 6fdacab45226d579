"""Contraction chains (ptc_tc_fksCuckoo_keyed): the output of one contraction
handed to the next with its key table.  Every chain's result is checked
against the oracle run on the oracle's own intermediate (coordinates
bit-exact after sorting, values within 1e-11 relative: the intermediate's
values already differ from the oracle's by up to 1e-12 relative and the next
contraction sums them); the table formed from the output's rows is checked
against a table built on the same output (same keys, same ids per key)."""
import numpy as np
import pytest

import partc_amd as P
from oracle import orc
from tests.helpers import max_rel, sort_coo

pytestmark = pytest.mark.gpu
TOL = 1e-11
METHODS = {"two_pass": P.METHOD_TWO_PASS, "upper_bound": P.METHOD_UPPER_BOUND, "learned": P.METHOD_LEARNED}


def tensor(seed, dims, nnz, dim_tc, nk):
    h, v = P.gen_coo(seed, dims, nnz, P.DIST_ZIPF, 1.1)
    return P.Tensor(h, v, None, np.array(dim_tc, np.uint32), nk)


def oracle_out(A, B, dim_tc, nk):
    """The oracle's contraction as a tensor with the next step's key split."""
    oh, ov, _ = orc.tc_fksCuckoo(A, B)
    return P.Tensor(oh, ov, None, np.array(dim_tc, np.uint32), nk)


def check(O, ref):
    h, v = sort_coo(O.hedges_array, O.val_array)
    rh, rv = sort_coo(ref.hedges_array, ref.val_array)
    assert h.shape == rh.shape, (h.shape, rh.shape)
    assert np.array_equal(h, rh)
    assert max_rel(v, rv) <= TOL


def same_table(ctx, HT, dO, key_modes):
    """HT (from the output's rows) against a table built on the same output:
    the same number of keys and, for every key, the same ids (a built table
    may order a slice's ids differently; the row table keeps output order)."""
    built = P.buildHashTable(dO, ctx)
    assert HT.info() == built.info()
    Oh = dO.download().hedges_array
    keys = np.unique(Oh[:, key_modes], axis=0)
    assert keys.shape[0] == HT.info()[0]
    a = HT.find(keys)
    b = built.find(keys)
    for x, y in zip(a, b):
        assert np.array_equal(x, np.sort(x))  # output order within a row
        assert np.array_equal(x, np.sort(y))
    built.free()


@pytest.fixture(scope="module")
def step1():
    """O(i, l) = sum_{j,k} A(i, j, k) B(j, k, l): A 300x40x50 free-first,
    B 40x50x200 contracted-first."""
    A = tensor(51, [300, 40, 50], 20000, [0, 1, 2], 1)
    B = tensor(52, [40, 50, 200], 20000, [0, 1, 2], 2)
    return A, B


@pytest.mark.parametrize("method", sorted(METHODS))
def test_chain_output_as_next_A(ctx, step1, method):
    """O keyed on its row mode i (next A side, free mode i, contracted l):
    the table comes from the output rows; O x C(l, m) matches the oracle."""
    if method == "learned" and not P.learned_available():
        pytest.skip("no trained row-size model in partc_amd/models")
    print(f"[case] chain O as next A, {method}", flush=True)
    A, B = step1
    C = tensor(53, [200, 150], 5000, [0, 1], 1)
    st = P.Stats()
    dO, HT_O, from_rows = P.tc_fksCuckoo_keyed(P.buildHashTable(A, ctx), P.buildHashTable(B, ctx), 1, [0, 1],
                                               stats=st, method=METHODS[method])
    assert from_rows
    print(f"  O nnz {st.nnz_out}, table {st.ms_build_out:.3f} ms")
    assert dO.info() == (2, st.nnz_out, 1)
    same_table(ctx, HT_O, dO, [0])
    O2 = P.tc_fksCuckoo(HT_O, P.buildHashTable(C, ctx))
    ref = orc.tc_fksCuckoo(oracle_out(A, B, [0, 1], 1), C)
    check(O2, P.Tensor(ref[0], ref[1], None, None, 2))


def test_chain_output_as_next_B(ctx, step1):
    """O keyed on its row mode i as the next contraction's B side:
    D(m, i) x O(i, l)."""
    print("[case] chain O as next B (row key)", flush=True)
    A, B = step1
    D = tensor(54, [120, 300], 6000, [0, 1], 1)  # free-first: free m, contracted i
    dO, HT_O, from_rows = P.tc_fksCuckoo_keyed(P.buildHashTable(A, ctx), P.buildHashTable(B, ctx), 1, [0, 1])
    assert from_rows
    same_table(ctx, HT_O, dO, [0])
    O3 = P.tc_fksCuckoo(P.buildHashTable(D, ctx), HT_O)
    ref = orc.tc_fksCuckoo(D, oracle_out(A, B, [0, 1], 1))
    check(O3, P.Tensor(ref[0], ref[1], None, None, 2))


def test_chain_column_key_is_built(ctx, step1):
    """A key on O's column mode l is not a row grouping: the table is built,
    and E(m, l) x O(l, i) still matches the oracle."""
    print("[case] chain column key", flush=True)
    A, B = step1
    E = tensor(55, [90, 200], 5000, [0, 1], 1)  # free m, contracted l
    dO, HT_O, from_rows = P.tc_fksCuckoo_keyed(P.buildHashTable(A, ctx), P.buildHashTable(B, ctx), 1, [1, 0])
    assert not from_rows
    O4 = P.tc_fksCuckoo(P.buildHashTable(E, ctx), HT_O)
    ref = orc.tc_fksCuckoo(E, oracle_out(A, B, [1, 0], 1))
    check(O4, P.Tensor(ref[0], ref[1], None, None, 2))


def test_chain_permuted_row_key(ctx):
    """Two row modes keyed in swapped order (j, i): the row keys are
    re-encoded.  O(i, j, m) = A(i, j, k, l) B(k, l, m), then
    O(j, i; m) x C(m, n)."""
    print("[case] chain permuted row key", flush=True)
    A = tensor(56, [30, 40, 25, 20], 15000, [0, 1, 2, 3], 2)
    B = tensor(57, [25, 20, 60], 8000, [0, 1, 2], 2)
    C = tensor(58, [60, 70], 3000, [0, 1], 1)
    dO, HT_O, from_rows = P.tc_fksCuckoo_keyed(P.buildHashTable(A, ctx), P.buildHashTable(B, ctx), 2, [1, 0, 2],
                                               method=P.METHOD_UPPER_BOUND)
    assert from_rows
    same_table(ctx, HT_O, dO, [1, 0])
    O2 = P.tc_fksCuckoo(HT_O, P.buildHashTable(C, ctx))
    ref = orc.tc_fksCuckoo(oracle_out(A, B, [1, 0, 2], 2), C)
    check(O2, P.Tensor(ref[0], ref[1], None, None, 3))


def test_chain_matches_built_table_nips_shape(ctx):
    """The nips-shape self-contraction's output (2482 x 2482 rows x rows, 6.1 M
    nonzeros) keyed on its rows: the row table equals a built one, and
    contracting the output with itself again gives the same result through
    either table."""
    print("[case] chain nips shape", flush=True)
    h, v = P.gen_coo(2, [2482, 2862, 14036, 17], 3101609, P.DIST_ZIPF, 1.1)
    A, B = P.self_contraction_inputs(h, v, [1, 2, 3])
    st = P.Stats()
    dO, HT_O, from_rows = P.tc_fksCuckoo_keyed(P.buildHashTable(A, ctx), P.buildHashTable(B, ctx), 1, [0, 1],
                                               stats=st, method=P.METHOD_UPPER_BOUND)
    assert from_rows
    print(f"  O nnz {st.nnz_out}, row table {st.ms_build_out:.3f} ms")
    built = P.buildHashTable(dO, ctx)
    assert HT_O.info() == built.info()
    # sum_l O(i, l) O(l, m): O is both sides, keyed on its row mode (A side:
    # free i, contracted l; B side: contracted i, free l)
    r1 = P.tc_fksCuckoo(HT_O, HT_O, method=P.METHOD_UPPER_BOUND)
    r2 = P.tc_fksCuckoo(built, built, method=P.METHOD_UPPER_BOUND)
    h1, v1 = sort_coo(r1.hedges_array, r1.val_array)
    h2, v2 = sort_coo(r2.hedges_array, r2.val_array)
    assert np.array_equal(h1, h2)
    assert max_rel(v1, v2) <= 1e-12
    built.free()


def test_chain_fp32(ctx, step1):
    """FP32 outputs: the row table shares the FP32 values; the next FP32
    contraction equals the one through a built table."""
    print("[case] chain fp32", flush=True)
    A, B = step1
    C = tensor(53, [200, 150], 5000, [0, 1], 1)
    dA, dB, dC = (P.DeviceTensor.upload_f32(T, ctx) for T in (A, B, C))
    dO, HT_O, from_rows = P.tc_fksCuckoo_keyed(P.buildHashTable(dA, ctx), P.buildHashTable(dB, ctx), 1, [0, 1],
                                               value_type=P.VAL_F32_F64ACC)
    assert from_rows
    HC = P.buildHashTable(dC, ctx)
    r1 = P.tc_fksCuckoo(HT_O, HC, download=False, value_type=P.VAL_F32_F64ACC)
    built = P.buildHashTable(dO, ctx)
    r2 = P.tc_fksCuckoo(built, HC, download=False, value_type=P.VAL_F32_F64ACC)
    t1, t2 = r1.download(), r2.download()
    h1, v1 = sort_coo(t1.hedges_array, t1.val_array)
    h2, v2 = sort_coo(t2.hedges_array, t2.val_array)
    assert np.array_equal(h1, h2)
    assert max_rel(v1, v2) <= 1e-6
    built.free()


def test_chain_errors(ctx, step1):
    A, B = step1
    HA, HB = P.buildHashTable(A, ctx), P.buildHashTable(B, ctx)
    for nk, perm in [(1, [0, 0]), (0, [0, 1]), (3, [0, 1]), (1, [0, 2])]:
        with pytest.raises(P.PtcError) as ei:
            P.tc_fksCuckoo_keyed(HA, HB, nk, perm)
        assert ei.value.code == -1
