"""The FP32 value path: FP32 tensors, products and sums compensated by
TwoProduct + TwoSum (PTC_VAL_F32_TWOSUM), or exact products summed in FP64
accumulators (PTC_VAL_F32_F64ACC); both give an FP32 (hi, lo) pair, checked
against

  * exact: the oracle on the same FP32 values widened to FP64 -- a product of
    two FP32 numbers is exact in FP64 and the oracle's FP64 sums are ~2^29
    times more precise than FP32, so this is the exact result of the FP32
    inputs for the purpose of these bounds;
  * the FP64 oracle on the original FP64 values (the error the FP32 path
    carries as a whole, input rounding included), reported per case.

Bars (written here): coordinates bit-exact; hi = fl32(sum) within one FP32
ulp of exact (relative 2^-23); hi + lo within 1e-9 relative of exact; hi
within 4e-7 relative of the FP64 oracle (two input roundings per product and
the final rounding, 3 * 2^-24, plus margin).  The uncompensated FP32 path
(PTC_VAL_F32_PLAIN) is run beside it to show what compensation buys.
"""
import numpy as np
import pytest

import partc_amd as P
from oracle import orc
from tests.helpers import golden_cases, max_rel, sort_coo, tc_inputs

pytestmark = pytest.mark.gpu

ULP_REL = 2.0 ** -23
PAIR_REL = 1e-9
VS_F64_REL = 4e-7
CASES = golden_cases()
ACCS = {"auto": P.ACC_AUTO, "hash": P.ACC_HASH, "spill": P.ACC_SPILL}
METHODS = {"two_pass": P.METHOD_TWO_PASS, "upper_bound": P.METHOD_UPPER_BOUND}
VTS = {"twosum": P.VAL_F32_TWOSUM, "f64acc": P.VAL_F32_F64ACC}


def with_vals(T, vals):
    return P.Tensor(T.hedges_array, vals, T.dimensions_array, T.dim_tc, T.num_key_dimensions)


def f32_of(T):
    return T.val_array.astype(np.float32)


def run_f32(ctx, A, B, value_type=P.VAL_F32_TWOSUM, **opts):
    dA, dB = P.DeviceTensor.upload_f32(A, ctx), P.DeviceTensor.upload_f32(B, ctx)
    HB, HA = P.buildHashTable(dB, ctx), P.buildHashTable(dA, ctx)
    st = P.Stats()
    O = P.tc_fksCuckoo(HA, HB, download=False, stats=st, value_type=value_type, **opts)
    h = O.download().hedges_array
    hi, lo = O.download_f32(with_lo=True)
    o = np.lexsort(h.T[::-1]) if h.shape[0] else np.arange(0)
    return h[o], hi[o], lo[o], st


def exact_f32_inputs(A, B):
    """Oracle on the FP32-rounded values (widened exactly)."""
    A2 = with_vals(A, f32_of(A).astype(np.float64))
    B2 = with_vals(B, f32_of(B).astype(np.float64))
    oh, ov, _ = orc.tc_fksCuckoo(A2, B2)
    return sort_coo(oh, ov)


def check_case(ctx, A, B, label, ref64=None, **opts):
    print(f"[case] f32 {label} opts={opts}", flush=True)
    h, hi, lo, st = run_f32(ctx, A, B, **opts)
    eh, ev = exact_f32_inputs(A, B)
    assert np.array_equal(h, eh)
    e_hi = max_rel(hi, ev)
    e_pair = max_rel(hi.astype(np.float64) + lo.astype(np.float64), ev)
    if ref64 is None:
        oh, ov, _ = orc.tc_fksCuckoo(A, B)
        ref64 = sort_coo(oh, ov)
    assert np.array_equal(h, ref64[0])
    e64 = max_rel(hi, ref64[1])
    print(f"  max rel err: hi vs exact {e_hi:.3e}, hi+lo vs exact {e_pair:.3e}, hi vs FP64 oracle {e64:.3e}")
    assert e_hi <= ULP_REL
    assert e_pair <= PAIR_REL
    assert e64 <= VS_F64_REL
    return st, (e_hi, e_pair, e64)


@pytest.mark.parametrize("vt", sorted(VTS))
@pytest.mark.parametrize("acc", sorted(ACCS))
@pytest.mark.parametrize("method", sorted(METHODS))
@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_f32(ctx, name, method, acc, vt):
    case = CASES[name]
    A, B = tc_inputs(case)
    st, _ = check_case(ctx, A, B, f"golden {name} {vt}", ref64=(case["coords"], case["vals_ordered"]),
                       method=METHODS[method], acc_mode=ACCS[acc], value_type=VTS[vt])
    assert st.nnz_out == case["coords"].shape[0]


@pytest.mark.parametrize("vt", sorted(VTS))
@pytest.mark.parametrize("chunk", [64, 1000])
@pytest.mark.parametrize("acc", sorted(ACCS))
def test_heavy_rows_f32(ctx, chunk, acc, vt):
    """Split rows: partial (sum, compensation) slabs merged in chunk order."""
    h, v = P.gen_coo(77, [60, 50, 40], 20000, P.DIST_ZIPF)
    A, B = P.self_contraction_inputs(h, v, [1, 2])
    st, _ = check_case(ctx, A, B, f"heavy chunk={chunk} {vt}", chunk_flops=chunk, acc_mode=ACCS[acc],
                       value_type=VTS[vt])
    assert st.heavy_rows > 0


def test_wide_columns_f32(ctx):
    """Column space too wide for dense LDS: LDS hash and HBM rows."""
    h, v = P.gen_coo(78, [300, 400, 5000], 6000, P.DIST_ZIPF)
    A, B = P.self_contraction_inputs(h, v, [0])
    for lds in (16, 4096):
        for m in METHODS.values():
            for vt in VTS.values():
                check_case(ctx, A, B, f"lds={lds} vt={vt}", lds_hash_max=lds, method=m, value_type=vt)


def test_unsafe_values_use_flags_f32(ctx):
    """Values whose products can underflow FP32 disable the -0.0 sentinel;
    results stay exact in coordinates."""
    h, v = P.gen_coo(79, [50, 40, 30], 8000, P.DIST_ZIPF)
    v = v.copy()
    v[::7] = 1e-30  # products of two such values underflow FP32
    A, B = P.self_contraction_inputs(h, v, [1, 2])
    print("[case] f32 tiny values", flush=True)
    hh, hi, lo, st = run_f32(ctx, A, B)
    eh, ev = exact_f32_inputs(A, B)
    assert np.array_equal(hh, eh)


def test_compensation_beats_plain_fp32(ctx):
    """Long sums: compensated FP32 keeps one-ulp accuracy; plain FP32 atomic
    accumulation does not."""
    h, v = P.gen_coo(80, [16, 40, 100], 50000, P.DIST_UNIFORM)  # ~2400 terms per output entry
    A, B = P.self_contraction_inputs(h, v, [1, 2])
    print("[case] f32 long sums: twosum vs plain", flush=True)
    eh, ev = exact_f32_inputs(A, B)
    h1, hi1, _, _ = run_f32(ctx, A, B, value_type=P.VAL_F32_TWOSUM)
    h2, hi2, lo2, _ = run_f32(ctx, A, B, value_type=P.VAL_F32_PLAIN)
    h3, hi3, _, _ = run_f32(ctx, A, B, value_type=P.VAL_F32_F64ACC)
    assert np.array_equal(h1, eh) and np.array_equal(h2, eh) and np.array_equal(h3, eh)
    assert np.all(lo2 == 0)
    e_comp, e_plain, e_wide = max_rel(hi1, ev), max_rel(hi2, ev), max_rel(hi3, ev)
    print(f"  max rel err vs exact: twosum {e_comp:.3e}, plain {e_plain:.3e}, f64acc {e_wide:.3e}")
    assert e_comp <= ULP_REL and e_wide <= ULP_REL
    assert e_plain > 4 * e_comp


@pytest.mark.parametrize("vt", sorted(VTS))
@pytest.mark.parametrize("scale", [0.01, 1.15])
def test_learned_f32_with_fallback(ctx, scale, vt):
    if not P.learned_available():
        pytest.skip("no trained row-size model in partc_amd/models")
    h, v = P.gen_coo(960, [300, 60, 80, 9], 40000, P.DIST_ZIPF)
    A, B = P.self_contraction_inputs(h, v, [1, 2, 3])
    st, _ = check_case(ctx, A, B, f"learned scale={scale} {vt}", method=P.METHOD_LEARNED, predict_scale=scale,
                       value_type=VTS[vt])
    if scale <= 0.01:
        assert st.fallback_rows > 0


def test_value_type_mismatches_are_errors(ctx):
    case = CASES["two_perm"]
    A, B = tc_inputs(case)
    d64A, d64B = P.DeviceTensor.upload(A, ctx), P.DeviceTensor.upload(B, ctx)
    d32A, d32B = P.DeviceTensor.upload_f32(A, ctx), P.DeviceTensor.upload_f32(B, ctx)
    H64A, H64B = P.buildHashTable(d64A, ctx), P.buildHashTable(d64B, ctx)
    H32A, H32B = P.buildHashTable(d32A, ctx), P.buildHashTable(d32B, ctx)
    for HA, HB, vt in ((H64A, H64B, P.VAL_F32_TWOSUM), (H32A, H32B, P.VAL_F64), (H32A, H64B, P.VAL_F32_TWOSUM),
                       (H64A, H64B, 7), (H32A, H32B, -1)):
        with pytest.raises(P.PtcError) as ei:
            P.tc_fksCuckoo(HA, HB, value_type=vt)
        assert ei.value.code == -1
    with pytest.raises(P.PtcError) as ei:
        P.tc_fksCuckoo(H64A, H64B, deterministic=1)
    assert ei.value.code == -6
