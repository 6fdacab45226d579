"""Edge cases of the boundary and of the accumulators' zero handling.

* Work items are capped at 2^31 multiply-adds (row passes keep 32-bit flop
  offsets inside an item): a larger chunk_flops is PTC_ERR_RANGE.
* A coordinate of 2^32-1 is PTC_ERR_RANGE at upload (key radices are the
  largest index + 1 in 32 bits).
* Sums that are exactly zero keep their coordinate, as the reference keeps
  every coordinate with a contributing pair (fksCuckooTC.hpp:318-330 adds
  without testing the sum): signed values that cancel, and zero values
  brought in by ptc_dtensor_set_values after a safe upload.
* An operand with no nonzero is PTC_ERR_EMPTY, also on tables used again
  once A's key count (0) is known.
"""
import numpy as np
import pytest

import partc_amd as P
from oracle import orc
from tests.helpers import max_rel, sort_coo

pytestmark = pytest.mark.gpu

METHODS = {"two_pass": P.METHOD_TWO_PASS, "upper_bound": P.METHOD_UPPER_BOUND, "learned": P.METHOD_LEARNED}
ACCS = {"auto": P.ACC_AUTO, "hash": P.ACC_HASH, "spill": P.ACC_SPILL}


def two_operands(seed, vals="uniform"):
    """A: 300 rows x 50 keys, B: 50 keys x 400 columns, 2000 nnz each:
    about 1.3 contributing pairs per output entry, so many entries hold one
    product and many hold two."""
    rng = np.random.default_rng(seed)
    ha, _ = P.gen_coo(seed, [300, 50], 2000, P.DIST_UNIFORM)
    hb, _ = P.gen_coo(seed + 1, [50, 400], 2000, P.DIST_UNIFORM)
    if vals == "sign":
        va = rng.choice([-1.0, 1.0], ha.shape[0])
        vb = rng.choice([-1.0, 1.0], hb.shape[0])
    else:
        va = rng.uniform(0.5, 1.0, ha.shape[0])
        vb = rng.uniform(0.5, 1.0, hb.shape[0])
    A = P.Tensor(ha, va, np.array([300, 50], np.uint32), np.array([0, 1], np.uint32), 1)  # free-first
    B = P.Tensor(hb, vb, np.array([50, 400], np.uint32), np.array([0, 1], np.uint32), 1)  # key first
    return A, B


def oracle(A, B):
    oh, ov, info = orc.tc_fksCuckoo(A, B)
    return sort_coo(oh, ov), info


def gpu(ctx, A, B, **opts):
    st = P.Stats()
    O = P.tc_fksCuckoo(P.buildHashTable(A, ctx), P.buildHashTable(B, ctx), stats=st, **opts)
    h, v = sort_coo(O.hedges_array, O.val_array)
    return h, v, st


def test_chunk_limit(ctx):
    h, v = P.gen_coo(1, [200, 300, 400], 20000, P.DIST_UNIFORM)
    A, B = P.self_contraction_inputs(h, v, [1, 2])
    print("[case] chunk_flops above 2^31", flush=True)
    with pytest.raises(P.PtcError) as e:
        gpu(ctx, A, B, chunk_flops=(1 << 31) + 1)
    assert e.value.code == -5
    print("[case] chunk_flops = 2^31", flush=True)
    hh, vv, _ = gpu(ctx, A, B, chunk_flops=1 << 31)
    (oh, ov), _ = oracle(A, B)
    assert np.array_equal(hh, oh)
    assert max_rel(vv, ov) <= 1e-12


def test_largest_index_rejected(ctx):
    h = np.array([[0, 1], [5, 0xFFFFFFFF]], dtype=np.uint32)
    t = P.Tensor(h, np.array([1.0, 2.0]), np.array([6, 0xFFFFFFFF], np.uint32), np.array([1, 0], np.uint32), 1)
    print("[case] coordinate 2^32-1 at upload", flush=True)
    with pytest.raises(P.PtcError) as e:
        P.DeviceTensor.upload(t, ctx)
    assert e.value.code == -5
    h[1, 1] = 0xFFFFFFFE  # the largest index is accepted
    t = P.Tensor(h, np.array([1.0, 2.0]), np.array([6, 0xFFFFFFFF], np.uint32), np.array([1, 0], np.uint32), 1)
    P.DeviceTensor.upload(t, ctx).free()


@pytest.mark.parametrize("acc", sorted(ACCS))
@pytest.mark.parametrize("method", sorted(METHODS))
def test_exact_cancellation_keeps_zero_sums(ctx, method, acc):
    if method == "learned" and not P.learned_available():
        pytest.skip("no trained row-size model in partc_amd/models")
    A, B = two_operands(21, vals="sign")
    (oh, ov), info = oracle(A, B)
    assert (ov == 0).sum() > 20  # the case is meaningful: many sums cancel exactly
    print(f"[case] +-1 values, {int((ov == 0).sum())} zero sums, {method} {acc}", flush=True)
    h, v, st = gpu(ctx, A, B, method=METHODS[method], acc_mode=ACCS[acc])
    assert st.flops == info["flops"]
    assert np.array_equal(h, oh)
    assert np.array_equal(v, ov)  # small integers: exact


@pytest.mark.parametrize("method", ["two_pass", "upper_bound"])
def test_zero_values_after_set_values(ctx, method):
    """Safe values at upload, then zeros (+0.0 and -0.0) written in place:
    the contraction must track touched entries with flags, or an entry whose
    only product is -0.0 would read as untouched and vanish."""
    A, B = two_operands(33)
    dA, dB = P.DeviceTensor.upload(A, ctx), P.DeviceTensor.upload(B, ctx)
    rng = np.random.default_rng(5)
    va = A.val_array.copy()
    zero = rng.random(va.shape[0]) < 0.2
    va[zero] = np.where(rng.random(zero.sum()) < 0.5, -0.0, 0.0)
    dA.set_values(va)
    A2 = P.Tensor(A.hedges_array, va, A.dimensions_array, A.dim_tc, A.num_key_dimensions)
    (oh, ov), info = oracle(A2, B)
    assert (ov == 0).sum() > 20
    print(f"[case] set_values with zeros, {int((ov == 0).sum())} zero sums, {method}", flush=True)
    st = P.Stats()
    O = P.tc_fksCuckoo(P.buildHashTable(dA, ctx), P.buildHashTable(dB, ctx), stats=st, method=METHODS[method])
    h, v = sort_coo(O.hedges_array, O.val_array)
    assert np.array_equal(h, oh)
    assert max_rel(v, ov) <= 1e-12


@pytest.mark.parametrize("method", ["two_pass", "upper_bound", "learned"])
def test_empty_operand_tables_reused(ctx, method):
    """A with no nonzero (and, separately, B with none): no pair contributes,
    so every call is PTC_ERR_EMPTY, as the reference's "tensor contraction
    not possible!!" (fksCuckooTC.hpp).  The same tables are contracted again
    after the first call and after the key count is read, when A's key
    count of 0 is known: the row passes then have no row to read."""
    if method == "learned" and not P.learned_available():
        pytest.skip("no trained row-size model")
    A, B = two_operands(35)
    empty_a = P.Tensor(np.zeros((0, 2), np.uint32), np.zeros(0), np.array([300, 50], np.uint32),
                       np.array([0, 1], np.uint32), 1)
    empty_b = P.Tensor(np.zeros((0, 2), np.uint32), np.zeros(0), np.array([50, 400], np.uint32),
                       np.array([0, 1], np.uint32), 1)
    for a, b, name in ((empty_a, B, "empty A"), (A, empty_b, "empty B")):
        print(f"[case] {name}, {method}", flush=True)
        HA, HB = P.buildHashTable(a, ctx), P.buildHashTable(b, ctx)
        for step in range(3):
            if step == 1:
                assert HA.info()[1] == a.N
            with pytest.raises(P.PtcError) as e:
                P.tc_fksCuckoo(HA, HB, method=METHODS[method])
            assert e.value.code == -4, (name, step)
    # the context still contracts after the empty calls
    (oh, ov), _ = oracle(A, B)
    O = P.tc_fksCuckoo(P.buildHashTable(A, ctx), P.buildHashTable(B, ctx), method=METHODS[method])
    h, v = sort_coo(O.hedges_array, O.val_array)
    assert np.array_equal(h, oh)
    assert max_rel(v, ov) <= 1e-12


def test_phase_timing_switch():
    """ptc_set_timing: off (the default) the phase times read 0 and the
    result is the one timing on gives, with the phase times filled."""
    A, B = two_operands(31)
    ctx = P.Context(0)
    try:
        dA, dB = P.DeviceTensor.upload(A, ctx), P.DeviceTensor.upload(B, ctx)
        outs = []
        for on in (False, True):
            ctx.set_timing(on)
            st = P.Stats()
            HB, HA = P.buildHashTable(dB, ctx), P.buildHashTable(dA, ctx)
            O = P.tc_fksCuckoo(HA, HB, stats=st, method=P.METHOD_UPPER_BOUND)
            outs.append(sort_coo(O.hedges_array, O.val_array))
            times = [st.ms_build_b, st.ms_build_a, st.ms_preprocess, st.ms_numeric, st.ms_kernel_numeric]
            if on:
                assert all(t > 0 for t in times), times
            else:
                assert all(t == 0 for t in times), times
        # coordinates bit-exact; sums within the FP64 tolerance (atomic
        # accumulation order differs between runs)
        assert np.array_equal(outs[0][0], outs[1][0])
        assert max_rel(outs[0][1], outs[1][1]) <= 1e-12
    finally:
        ctx.close()
