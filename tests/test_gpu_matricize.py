"""ptc_matricize_spGEMM as a second device method: the column-wise product
of cs_multiply (cs_multiply.c:6-38; matricizedspGEMM.hpp:172-603), B grouped
by its free modes (the output columns) and A by its contracted modes,
against the row-wise ptc_contract and the oracle -- the cross-check of the
two methods that the reference's disabled main.cpp:131-158 intended.

Bar: coordinates bit-exact after sorting, values within 1e-12 relative, for
both methods against the oracle and against each other; the column method's
output comes column by column (each output column one contiguous run).
"""
import numpy as np
import pytest

import partc_amd as P
from oracle import orc
from tests.helpers import golden_cases, max_rel, meta_tensors, sort_coo

pytestmark = pytest.mark.gpu
TOL = 1e-12
CASES = golden_cases()


def both_methods(ctx, A, B):
    st_c, st_m = P.Stats(), P.Stats()
    Oc = P.contract(A, B, ctx, stats=st_c)
    Om = P.matricize_spGEMM(A, B, ctx, stats=st_m)
    return Oc, Om, st_c, st_m


def columns_contiguous(O, ncolmodes):
    """Each output column (the last ncolmodes modes) is one contiguous run."""
    h = O.hedges_array
    if h.shape[0] < 2:
        return True
    cols = h[:, h.shape[1] - ncolmodes:]
    change = np.any(cols[1:] != cols[:-1], axis=1)
    runs = int(change.sum()) + 1
    distinct = np.unique(cols, axis=0).shape[0]
    return runs == distinct


def check_pair(ctx, A, B, label, ref=None):
    print(f"[case] matricized vs row-wise: {label}", flush=True)
    Oc, Om, st_c, st_m = both_methods(ctx, A, B)
    hc, vc = sort_coo(Oc.hedges_array, Oc.val_array)
    hm, vm = sort_coo(Om.hedges_array, Om.val_array)
    if ref is None:
        oh, ov, info = orc.contract_meta(A, B)
        ref = sort_coo(oh, ov)
        assert st_c.flops == info["flops"] and st_m.flops == info["flops"]
    assert np.array_equal(hm, hc)
    assert max_rel(vm, vc) <= TOL
    assert np.array_equal(hc, ref[0]) and np.array_equal(hm, ref[0])
    assert max_rel(vc, ref[1]) <= TOL and max_rel(vm, ref[1]) <= TOL
    assert st_m.nnz_a == A.N and st_m.nnz_b == B.N
    assert columns_contiguous(Om, B.d - int(B.num_key_dimensions))
    return Om, st_m


@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_both_methods(ctx, name):
    case = CASES[name]
    A, B = meta_tensors(case)
    check_pair(ctx, A, B, f"golden {name}", ref=(case["coords"], case["vals_dense"]))


def test_c1_both_methods(ctx):
    """3-mode 200x300x400, 50k uniform nnz, self-contraction over {1,2}."""
    h, v = P.gen_coo(1, [200, 300, 400], 50000, P.DIST_UNIFORM)
    T = P.Tensor(h, v, None, np.array([1, 2, 0], np.uint32), 2)
    check_pair(ctx, T, T, "C1")


TWO = [
    # A dims, A perm (contracted first), B dims, B perm, c, nnz of A, nnz of B, zipf
    ([300, 50, 20], [1, 2, 0], [50, 20, 400], [0, 1, 2], 2, 20000, 20000, False),
    ([300, 50, 20], [2, 1, 0], [400, 20, 50], [2, 1, 0], 2, 20000, 20000, True),  # modes paired out of order
    ([40, 30, 12000], [1, 0, 2], [30, 2000], [0, 1], 1, 8000, 2000, True),      # wide rows: the column method hashes
    ([2000, 60], [1, 0], [60, 20000], [0, 1], 1, 6000, 8000, True),            # wide columns: the row method hashes
    ([25, 20, 30, 15], [1, 2, 0, 3], [20, 30, 40], [0, 1, 2], 2, 30000, 20000, True),
]


@pytest.mark.parametrize("i", range(len(TWO)))
def test_two_operands_both_methods(ctx, i):
    da, pa, db, pb, c, nnz_a, nnz_b, zipf = TWO[i]
    dist = P.DIST_ZIPF if zipf else P.DIST_UNIFORM
    ha, va = P.gen_coo(100 + i, da, nnz_a, dist)
    hb, vb = P.gen_coo(200 + i, db, nnz_b, dist)
    A = P.Tensor(ha, va, None, np.array(pa, np.uint32), c)
    B = P.Tensor(hb, vb, None, np.array(pb, np.uint32), c)
    check_pair(ctx, A, B, f"two operands {da} x {db}")


def test_empty_both_methods(ctx):
    A = P.Tensor(np.array([[0, 1], [1, 2]], np.uint32), np.array([1.0, 2.0]), None, np.array([1, 0], np.uint32), 1)
    B = P.Tensor(np.array([[5, 1], [6, 2]], np.uint32), np.array([1.0, 2.0]), None, np.array([0, 1], np.uint32), 1)
    for fn in (P.contract, P.matricize_spGEMM):
        with pytest.raises(P.PtcError) as e:
            fn(A, B, ctx)
        assert e.value.code == -4
