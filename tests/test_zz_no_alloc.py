"""ptc_tc_fksCuckoo_noalloc: the allocation-free variant of ptc_tc_fksCuckoo
(scratch in one static device array per device, output in caller arrays).

Checked through the C ABI against ptc_tc_fksCuckoo (exact two-pass, FP64):
coordinates bit-exact after sorting, values within 1e-12 relative, on two
calls in a row -- a dense-accumulator contraction, then a smaller one with
hash accumulators that reuses the scratch the first left behind -- plus the
two size errors: output arrays too small (the size is still reported) and a
contraction whose scratch exceeds the array.
"""
import ctypes as C

import numpy as np
import pytest

import partc_amd as P
from tests import hiprt
from tests.helpers import max_rel, sort_coo

pytestmark = pytest.mark.gpu
TOL = 1e-12


def reference(ctx, HA, HB):
    O = P.tc_fksCuckoo(HA, HB, method=P.METHOD_TWO_PASS)
    return sort_coo(O.hedges_array, O.val_array)


def noalloc(ctx, HA, HB, cap, d_out):
    """One call with output arrays of cap entries; returns (rc, nnz, the
    coordinates and values written, stats)."""
    hed = hiprt.DeviceBuffer(4 * d_out * cap)
    val = hiprt.DeviceBuffer(8 * cap)
    nnz = C.c_uint64(0)
    st = P.Stats()
    rc = P.lib().ptc_tc_fksCuckoo_noalloc(ctx.handle, HA._h, HB._h, HA.tensor._h, HB.tensor._h, None,
                                          hed.ptr, val.ptr, cap, C.byref(nnz), C.byref(st))
    n = int(nnz.value)
    h = v = None
    if rc == 0:
        h = hed.to_numpy(np.uint32, n * d_out).reshape(n, d_out)
        v = val.to_numpy(np.float64, n)
    hed.free()
    val.free()
    return rc, n, h, v, st


def tables(ctx, h, v, contracted):
    A, B = P.self_contraction_inputs(h, v, contracted)
    return P.buildHashTable(A, ctx), P.buildHashTable(B, ctx)


def test_noalloc_matches_tc_on_two_calls(ctx):
    cases = [
        ("dense, 3-mode 200x300x400, 50k nnz, contract {1,2}", P.gen_coo(1, [200, 300, 400], 50000, P.DIST_UNIFORM),
         [1, 2], 2),
        ("hash rows (74 in HBM), 3-mode 12000x30x40, 8k Zipf nnz, contract {1,2}",
         P.gen_coo(2, [12000, 30, 40], 8000, P.DIST_ZIPF), [1, 2], 2),
    ]
    for label, (h, v), contracted, d_out in cases:
        print(f"[case] noalloc {label}", flush=True)
        HA, HB = tables(ctx, h, v, contracted)
        rh, rv = reference(ctx, HA, HB)
        rc, nnz, gh, gv, st = noalloc(ctx, HA, HB, rh.shape[0] + 7, d_out)
        assert rc == 0, ctx.last_error()
        assert nnz == rh.shape[0] and st.nnz_out == nnz
        gh, gv = sort_coo(gh, gv)
        assert np.array_equal(gh, rh)
        assert max_rel(gv, rv) <= TOL
        HA.free()
        HB.free()


def test_noalloc_size_errors(ctx):
    h, v = P.gen_coo(3, [200, 300, 400], 20000, P.DIST_UNIFORM)
    HA, HB = tables(ctx, h, v, [1, 2])
    rh, rv = reference(ctx, HA, HB)
    print("[case] noalloc output arrays one entry short", flush=True)
    rc, nnz, _, _, _ = noalloc(ctx, HA, HB, rh.shape[0] - 1, 2)
    assert rc == -5 and nnz == rh.shape[0]
    # the nips-shape output (6,128,310 nonzeros) does not fit one entry: its
    # size is still reported, from the symbolic pass
    print("[case] noalloc size query at nips shape", flush=True)
    h2, v2 = P.gen_coo(2, [2482, 2862, 14036, 17], 3101609, P.DIST_ZIPF, 1.1)
    HA2, HB2 = tables(ctx, h2, v2, [1, 2, 3])
    rc, nnz, _, _, _ = noalloc(ctx, HA2, HB2, 1, 2)
    assert rc == -5 and nnz == 6128310
    # 560 rows with HBM accumulators sized by their flop bounds need more
    # scratch than the array holds
    print("[case] noalloc scratch too small (12000 hash columns, 20k Zipf nnz)", flush=True)
    h3, v3 = P.gen_coo(2, [12000, 30, 40], 20000, P.DIST_ZIPF)
    HA3, HB3 = tables(ctx, h3, v3, [1, 2])
    r3h, _ = reference(ctx, HA3, HB3)
    rc, nnz, _, _, _ = noalloc(ctx, HA3, HB3, r3h.shape[0], 2)
    assert rc == -5
    assert "scratch" in ctx.last_error()
    # and the context still works afterwards, with and without the scratch
    rc, nnz, gh, gv, _ = noalloc(ctx, HA, HB, rh.shape[0], 2)
    assert rc == 0 and nnz == rh.shape[0]
    gh, gv = sort_coo(gh, gv)
    assert np.array_equal(gh, rh) and max_rel(gv, rv) <= TOL
    O = P.tc_fksCuckoo(HA2, HB2, download=False)
    assert O.info()[1] == 6128310
