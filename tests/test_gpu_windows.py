"""Dense column windows against the oracle.

A 16-bit column space too wide for one dense LDS accumulator is held a
window at a time: every window's workgroups stream a row's multiply-adds and
keep the columns in their range, and rows of at most 1024 multiply-adds go
to small LDS hash tables.  PTC_ACC_WINDOW forces the path (dense_max_cols
caps the window width, so small column spaces get several windows); AUTO
takes it for column spaces wider than one dense accumulator.

Bar: output coordinates bit-exact after sorting; FP64 values within 1e-12
relative (atomic summation order differs from the oracle's); FP32 modes
within the bounds of the FP32 tests.  The exact two-pass output must also
keep one dense accumulator's row layout (each row's entries contiguous, rows
in table order), since chain steps build the next key table from its rows.
"""
import numpy as np
import pytest

import partc_amd as P
from oracle import orc
from tests.helpers import max_rel, sort_coo

pytestmark = pytest.mark.gpu

TOL = 1e-12
METHODS = {"two_pass": P.METHOD_TWO_PASS, "upper_bound": P.METHOD_UPPER_BOUND, "learned": P.METHOD_LEARNED}

_ORACLE = {}


def case_nips_like():
    """4-mode Zipf tensor contracted over {1, 2}: 3600 x 3600 output, heavy
    and small rows."""
    h, v = P.gen_coo(91, [300, 40, 30, 12], 60000, P.DIST_ZIPF, 1.1)
    return P.self_contraction_inputs(h, v, [1, 2])


def case_wide():
    """Column space of 3000 x 12 = 36000 (wider than one dense accumulator,
    at most 4 windows): AUTO windows."""
    h, v = P.gen_coo(92, [3000, 60, 50, 12], 40000, P.DIST_ZIPF, 1.1)
    return P.self_contraction_inputs(h, v, [1, 2])


CASES = {"nips_like": case_nips_like, "wide": case_wide}


def oracle(name):
    """(A, B, oracle coordinates and values sorted, oracle info), once per case."""
    if name not in _ORACLE:
        A, B = CASES[name]()
        oh, ov, info = orc.tc_fksCuckoo(A, B)
        oh, ov = sort_coo(oh, ov)
        _ORACLE[name] = (A, B, oh, ov, info)
    return _ORACLE[name]


def run(ctx, A, B, **opts):
    if opts.get("method") == P.METHOD_LEARNED and not P.learned_available():
        pytest.skip("no trained row-size model")
    HB = P.buildHashTable(B, ctx)
    HA = P.buildHashTable(A, ctx)
    st = P.Stats()
    O = P.tc_fksCuckoo(HA, HB, stats=st, **opts)
    return O.hedges_array, O.val_array, st


def check(ctx, name, label, **opts):
    A, B, oh, ov, info = oracle(name)
    print(f"[case] windows {name} {label} opts={opts}", flush=True)
    h, v, st = run(ctx, A, B, **opts)
    hs, vs = sort_coo(h, v)
    assert st.flops == info["flops"]
    assert hs.shape == oh.shape, (hs.shape, oh.shape)
    assert np.array_equal(hs, oh)
    assert max_rel(vs, ov) <= TOL
    return h, v, st


def same_rows_as_dense(ctx, A, B, **wopts):
    """The windowed exact output has one dense accumulator's row layout (same
    tables, so the same row order): every row's entries contiguous, rows in
    the same order with the same counts.  Within a row, windowed rows list
    their columns ascending and small rows in hash-table order."""
    HB = P.buildHashTable(B, ctx)
    HA = P.buildHashTable(A, ctx)
    st = P.Stats()
    Od = P.tc_fksCuckoo(HA, HB, acc_mode=P.ACC_DENSE)
    Ow = P.tc_fksCuckoo(HA, HB, acc_mode=P.ACC_WINDOW, stats=st, **wopts)
    assert st.col_windows >= 1
    nr = 2  # row modes of the output
    assert np.array_equal(Od.hedges_array[:, :nr], Ow.hedges_array[:, :nr]), "row layout differs from the dense one"


@pytest.mark.parametrize("chunk", [0, 2000])
@pytest.mark.parametrize("method", sorted(METHODS))
def test_forced_windows(ctx, method, chunk):
    """Windows of at most 1024 columns over a 3600-column space (4 windows);
    chunk 2000 splits rows over many workgroups (per-window slabs, merges)."""
    opts = dict(method=METHODS[method], acc_mode=P.ACC_WINDOW, dense_max_cols=1024)
    if chunk:
        opts["chunk_flops"] = chunk
    h, v, st = check(ctx, "nips_like", f"{method} chunk={chunk}", **opts)
    assert st.col_windows == 4
    assert st.small_rows > 0
    if chunk:
        assert st.heavy_rows > 0
    if method != "two_pass":
        assert st.fallback_rows == 0


@pytest.mark.parametrize("width", [64, 960, 4096])
def test_window_widths(ctx, width):
    """One to four windows, the narrowest 64 columns (a single bitmap word
    pair per window), on the exact path."""
    cols = 3600
    nwin = -(-cols // width)
    if nwin > 4:
        with pytest.raises(P.PtcError):
            run(ctx, *oracle("nips_like")[:2], acc_mode=P.ACC_WINDOW, dense_max_cols=width)
        return
    h, v, st = check(ctx, "nips_like", f"width={width}", acc_mode=P.ACC_WINDOW, dense_max_cols=width)
    assert st.col_windows == nwin
    same_rows_as_dense(ctx, *oracle("nips_like")[:2], dense_max_cols=width)


@pytest.mark.parametrize("method", sorted(METHODS))
def test_auto_windows_wide(ctx, method):
    """AUTO on a 36000-column space picks windows instead of hash tables."""
    h, v, st = check(ctx, "wide", method, method=METHODS[method])
    assert st.col_windows >= 2


def test_hash_mode_still_hash(ctx):
    """PTC_ACC_HASH keeps hash tables on the same wide space."""
    h, v, st = check(ctx, "wide", "hash", acc_mode=P.ACC_HASH)
    assert st.col_windows == 0


def test_learned_outgrows_capacities(ctx):
    """A learned single pass whose capacities sum below the result (scale
    0.05) is redone on the exact path: same output, fallback counted."""
    h, v, st = check(ctx, "wide", "learned scale=0.05", method=P.METHOD_LEARNED, predict_scale=0.05)
    assert st.fallback_rows > 0
    assert st.col_windows >= 2


def test_windows_with_zero_products(ctx):
    """Values that cancel or are zero (touched bits tracked): every touched
    entry stays in the output, as with one dense accumulator."""
    A, B, oh, ov, info = oracle("nips_like")
    vals = A.val_array.copy()
    vals[::7] = 0.0
    vals[1::5] *= -1.0
    # A and B are two key splits of the same nonzeros, in the same order
    assert np.array_equal(A.hedges_array, B.hedges_array)
    A2 = P.Tensor(A.hedges_array, vals, A.dimensions_array, A.dim_tc, A.num_key_dimensions)
    B2 = P.Tensor(B.hedges_array, vals.copy(), B.dimensions_array, B.dim_tc, B.num_key_dimensions)
    oh2, ov2, _ = orc.tc_fksCuckoo(A2, B2)
    ohs, ovs = sort_coo(oh2, ov2)
    # sums of the terms' magnitudes bound each entry's rounding error
    Aa = P.Tensor(A.hedges_array, np.abs(vals), A.dimensions_array, A.dim_tc, A.num_key_dimensions)
    Ba = P.Tensor(B.hedges_array, np.abs(vals), B.dimensions_array, B.dim_tc, B.num_key_dimensions)
    ah, av, _ = orc.tc_fksCuckoo(Aa, Ba)
    ahs, avs = sort_coo(ah, av)
    assert np.array_equal(ahs, ohs)
    assert np.any(avs == 0.0)  # entries touched by zero products only
    for m in (P.METHOD_TWO_PASS, P.METHOD_UPPER_BOUND):
        print(f"[case] windows zero products method={m}", flush=True)
        h, v, st = run(ctx, A2, B2, method=m, acc_mode=P.ACC_WINDOW, dense_max_cols=1024)
        hs, vs = sort_coo(h, v)
        assert np.array_equal(hs, ohs)
        assert np.all(np.abs(vs - ovs) <= TOL * avs)


@pytest.mark.parametrize("vt", ["twosum", "f64acc"])
@pytest.mark.parametrize("method", ["two_pass", "upper_bound"])
def test_windows_f32(ctx, method, vt):
    """FP32 value types over windows, to the FP32 path's own bars (hi within
    one FP32 ulp of the exact sum of the FP32 inputs, hi + lo within 1e-9)."""
    from tests.test_gpu_f32 import check_case

    A, B, oh, ov, info = oracle("nips_like")
    vts = {"twosum": P.VAL_F32_TWOSUM, "f64acc": P.VAL_F32_F64ACC}
    st, _ = check_case(ctx, A, B, f"windows {method} {vt}", ref64=(oh, ov), method=METHODS[method],
                       value_type=vts[vt], acc_mode=P.ACC_WINDOW, dense_max_cols=1024)
    assert st.col_windows == 4
