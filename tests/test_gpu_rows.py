"""Row preprocessing at the boundaries of its one-workgroup form, against
the oracle.

Up to 8191 dense rows take one workgroup for their flops, the two row
scans, the split-row list and the upper-bound capacities (four rows a
thread up to 4095 rows, eight up to 8191); more rows take the multi-launch
form.  The work items then find their rows in the scanned counts.  The
learned method's capacity kernel reads each row's hit count and longest
slice from the positions of the 1024-position parts the row only partly
covers and from per-part records for the parts inside it: rows here cover
one part, straddle two, and span many.  Every case runs each method and
compares with the oracle (coordinates bit-exact after sorting, values
within 1e-12 relative) and the flop count.

Key tables used again: the first contraction reads A's key count back, and
later ones size every row array by that count instead of by the bound on
distinct keys (the row passes then read no count from the device).
"""
import numpy as np
import pytest

import partc_amd as P
from oracle import orc
from tests.helpers import max_rel, sort_coo

pytestmark = pytest.mark.gpu
TOL = 1e-12
METHODS = {"two_pass": P.METHOD_TWO_PASS, "upper_bound": P.METHOD_UPPER_BOUND, "learned": P.METHOD_LEARNED}


def matrices(nrows, row_len, nkeys, ncols, slice_len, seed):
    """A(i, j): row i holds row_len(i) distinct keys j < nkeys; B(j, k): key
    j holds slice_len distinct columns k < ncols (a dense output space)."""
    rng = np.random.default_rng(seed)
    ah, av = [], []
    for i in range(nrows):
        n = min(nkeys, row_len(i))
        for j in rng.choice(nkeys, size=n, replace=False):
            ah.append((i, int(j)))
        av += list(rng.uniform(0.5, 1.0, size=n))
    bh = [(j, int(k)) for j in range(nkeys) for k in rng.choice(ncols, size=slice_len, replace=False)]
    bv = rng.uniform(0.5, 1.0, size=len(bh))
    A = P.Tensor(np.array(ah, np.uint32), np.array(av), None, np.array([0, 1], np.uint32), 1)
    B = P.Tensor(np.array(bh, np.uint32), bv, None, np.array([0, 1], np.uint32), 1)
    return A, B


def check(ctx, A, B, label):
    oh, ov, info = orc.tc_fksCuckoo(A, B)
    oh, ov = sort_coo(oh, ov)
    for name, m in METHODS.items():
        if name == "learned" and not P.learned_available():  # loads the trained model on first use
            pytest.skip("no trained row-size model in partc_amd/models")
        print(f"[case] rows {label} {name}", flush=True)
        HB, HA = P.buildHashTable(B, ctx), P.buildHashTable(A, ctx)
        st = P.Stats()
        O = P.tc_fksCuckoo(HA, HB, stats=st, method=m)
        h, v = sort_coo(O.hedges_array, O.val_array)
        assert st.flops == info["flops"], name
        assert h.shape == oh.shape and np.array_equal(h, oh), name
        assert max_rel(v, ov) <= TOL, name


@pytest.mark.parametrize("nrows", [1, 4095, 4096, 8191, 8192, 9000])
def test_row_counts_around_one_workgroup(ctx, nrows):
    """Rows either side of the one-workgroup limits (4095 / 8191) and past
    them; every row short (one to three keys)."""
    A, B = matrices(nrows, lambda i: 1 + i % 3, nkeys=400, ncols=300, slice_len=3, seed=nrows)
    check(ctx, A, B, f"{nrows} short rows")


def test_rows_spanning_many_parts(ctx):
    """A few rows of 5000-20000 positions (many 1024-position parts inside
    one row, rows starting and ending inside parts) beside short rows, and
    heavy rows split into several work items."""
    lens = [20000, 1, 5000, 3, 1023, 1025, 7000, 2, 2048, 12000]
    A, B = matrices(len(lens), lambda i: lens[i], nkeys=20000, ncols=2000, slice_len=4, seed=7)
    check(ctx, A, B, "long rows")


@pytest.mark.parametrize("nrows,ncols", [(4095, 300), (9000, 300), (3000, 200000)])
def test_tables_used_again(ctx, nrows, ncols):
    """The same two key tables contracted by every method in turn, twice:
    dense rows in one workgroup (4095) and in the multi-launch form (9000),
    and hash-accumulated rows (200000 columns)."""
    A, B = matrices(nrows, lambda i: 1 + i % 3, nkeys=400, ncols=ncols, slice_len=3, seed=nrows + 1)
    oh, ov, info = orc.tc_fksCuckoo(A, B)
    oh, ov = sort_coo(oh, ov)
    names = [n for n in METHODS if n != "learned" or P.learned_available()]
    HB, HA = P.buildHashTable(B, ctx), P.buildHashTable(A, ctx)
    for rep in range(2):
        for name in names:
            print(f"[case] tables used again, {nrows} rows x {ncols} columns, call {rep} {name}", flush=True)
            st = P.Stats()
            O = P.tc_fksCuckoo(HA, HB, stats=st, method=METHODS[name])
            h, v = sort_coo(O.hedges_array, O.val_array)
            assert st.flops == info["flops"], name
            assert h.shape == oh.shape and np.array_equal(h, oh), name
            assert max_rel(v, ov) <= TOL, name
