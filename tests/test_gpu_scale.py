"""Parity at the full sizes of the benchmark configurations.

* nips-shape 4-mode self-contraction (3,101,609 nnz, F = 1.39e8): the whole
  output against the oracle -- coordinates bit-exact, values within 1e-12.
* two-tensor contraction (A 10000x5000x200, B 5000x200x8000, 5M nnz each,
  F ~ 2.6e9): too large for the oracle as a whole, so
    - 40 sampled output rows against the oracle run on A restricted to those
      rows (all of B);
    - every output row's value sum against the closed form
      sum_s O[i,s] = sum_{a in row i} A[a] * SB(key(a)), with SB(k) the sum of
      B's values with contracted key k (linearity; numpy, FP64);
    - every output row's column-weighted sum sum_s O[i,s] w(s) against the
      same closed form over B's values times w(column), with w a fixed
      pseudo-random weight in [1, 2) per column: a column moved to another
      place inside a row changes it;
    - the output count against the learned and upper-bound methods.
* the FP32 value path (FP32 inputs: the values above rounded to FP32) on
  both tensors, each value mode: coordinates bit-exact; against the exact
  result of the FP32 inputs (the oracle on them, widened to FP64), hi within
  one FP32 ulp (2^-23 relative) for the compensated modes, and within the
  recursive-summation bound (n + 1) 2^-24 for plain FP32 sums of n positive
  products (n bounded per entry by the contracted keys the two rows share);
  the error against the FP64 oracle on the original values is printed per
  result.
"""
import numpy as np
import pytest

import partc_amd as P
from oracle import orc
from tests.helpers import max_rel, sort_coo

pytestmark = pytest.mark.gpu
TOL = 1e-12
METHODS = {"two_pass": P.METHOD_TWO_PASS, "upper_bound": P.METHOD_UPPER_BOUND, "learned": P.METHOD_LEARNED}


def gpu_contract(ctx, A, B, **opts):
    st = P.Stats()
    O = P.tc_fksCuckoo(P.buildHashTable(A, ctx), P.buildHashTable(B, ctx), stats=st, **opts)
    return O, st


@pytest.fixture(scope="module")
def nips():
    h, v = P.gen_coo(2, [2482, 2862, 14036, 17], 3101609, P.DIST_ZIPF, 1.1)
    A, B = P.self_contraction_inputs(h, v, [1, 2, 3])
    oh, ov, info = orc.tc_fksCuckoo(A, B)
    return A, B, sort_coo(oh, ov), info


@pytest.mark.parametrize("method", sorted(METHODS))
def test_nips_full_output(ctx, nips, method):
    if method == "learned" and not P.learned_available():
        pytest.skip("no trained row-size model in partc_amd/models")
    A, B, (oh, ov), info = nips
    print(f"[case] nips full size {method}", flush=True)
    O, st = gpu_contract(ctx, A, B, method=METHODS[method])
    h, v = sort_coo(O.hedges_array, O.val_array)
    assert st.flops == info["flops"]
    assert np.array_equal(h, oh)
    assert max_rel(v, ov) <= TOL
    print(f"  nnz {h.shape[0]} fallback rows {st.fallback_rows}")


def test_nips_matricized_method(ctx, nips):
    """The column-wise matricized product (ptc_matricize_spGEMM) on the
    nips-shape tensor against the oracle's full output."""
    A, B, (oh, ov), info = nips
    print("[case] nips full size, column-wise matricized product", flush=True)
    Am = P.Tensor(A.hedges_array, A.val_array, A.dimensions_array, B.dim_tc, B.num_key_dimensions)
    st = P.Stats()
    O = P.matricize_spGEMM(Am, B, ctx, stats=st)
    h, v = sort_coo(O.hedges_array, O.val_array)
    assert st.flops == info["flops"]
    assert np.array_equal(h, oh)
    assert max_rel(v, ov) <= TOL


def two_tensor():
    ha, va = P.gen_coo(3, [10000, 5000, 200], 5_000_000, P.DIST_ZIPF, 1.1)
    hb, vb = P.gen_coo(4, [5000, 200, 8000], 5_000_000, P.DIST_ZIPF, 1.1)
    A = P.make_free_first(P.Tensor(ha, va, None, np.array([1, 2, 0], np.uint32), 2))
    B = P.Tensor(hb, vb, None, np.array([0, 1, 2], np.uint32), 2)
    return A, B


def row_sums_closed_form(A, B):
    """sum_s O[i, s] for every A row i, from sums of B over each key."""
    hb, vb = B.hedges_array, B.val_array
    kb = hb[:, 0].astype(np.int64) * 200 + hb[:, 1]
    SB = np.bincount(kb, weights=vb, minlength=5000 * 200)
    ha, va = A.hedges_array, A.val_array
    ka = ha[:, 1].astype(np.int64) * 200 + ha[:, 2]
    ok = ka < SB.shape[0]
    return np.bincount(ha[ok, 0], weights=va[ok] * SB[ka[ok]], minlength=10000)


@pytest.fixture(scope="module")
def c3():
    A, B = two_tensor()
    rs = row_sums_closed_form(A, B)
    rng = np.random.default_rng(11)
    present = np.unique(A.hedges_array[:, 0])
    rows = np.sort(rng.choice(present[present >= 100], 40, replace=False))
    sel = np.isin(A.hedges_array[:, 0], rows)
    Asub = P.Tensor(A.hedges_array[sel], A.val_array[sel], A.dimensions_array, A.dim_tc, A.num_key_dimensions)
    oh, ov, _ = orc.tc_fksCuckoo(Asub, B)
    return A, B, rs, rows, sort_coo(oh, ov)


@pytest.mark.parametrize("method", sorted(METHODS))
def test_two_tensor_full_size(ctx, c3, method):
    if method == "learned" and not P.learned_available():
        pytest.skip("no trained row-size model in partc_amd/models")
    A, B, rs, rows, (oh, ov) = c3
    print(f"[case] two-tensor full size {method}", flush=True)
    O, st = gpu_contract(ctx, A, B, method=METHODS[method])
    h, v = O.hedges_array, O.val_array
    print(f"  F {st.flops} nnz {st.nnz_out} fallback rows {st.fallback_rows} of {st.rows}")
    # coordinates unique, inside the output space
    assert h.shape[1] == 2 and h[:, 0].max() < 10000 and h[:, 1].max() < 8000
    key = h[:, 0].astype(np.int64) * 8000 + h[:, 1]
    assert np.unique(key).shape[0] == key.shape[0]
    # every row's value sum against the closed form
    gs = np.bincount(h[:, 0], weights=v, minlength=10000)
    nz = rs > 0
    assert np.array_equal(gs > 0, nz)
    assert max_rel(gs[nz], rs[nz]) <= 1e-10
    # sampled rows exactly
    sel = np.isin(h[:, 0], rows)
    hs, vs = sort_coo(h[sel], v[sel])
    assert np.array_equal(hs, oh)
    assert max_rel(vs, ov) <= TOL


def column_weight(s):
    """w(s) in [1, 2): a fixed pseudo-random weight of output column s."""
    s = np.asarray(s, dtype=np.uint64)
    return 1.0 + ((s * np.uint64(2654435761)) % np.uint64(1 << 32)).astype(np.float64) / 2.0 ** 32


def weighted_row_sums_closed_form(A, B):
    """sum_s O[i, s] w(s) for every A row i, from B's values times w(col)."""
    hb, vb = B.hedges_array, B.val_array
    kb = hb[:, 0].astype(np.int64) * 200 + hb[:, 1]
    SBw = np.bincount(kb, weights=vb * column_weight(hb[:, 2]), minlength=5000 * 200)
    ha, va = A.hedges_array, A.val_array
    ka = ha[:, 1].astype(np.int64) * 200 + ha[:, 2]
    ok = ka < SBw.shape[0]
    return np.bincount(ha[ok, 0], weights=va[ok] * SBw[ka[ok]], minlength=10000)


@pytest.mark.parametrize("method", ["two_pass", "learned"])
def test_two_tensor_column_checksum(ctx, c3, method):
    if method == "learned" and not P.learned_available():
        pytest.skip("no trained row-size model in partc_amd/models")
    A, B = c3[0], c3[1]
    ws = weighted_row_sums_closed_form(A, B)
    print(f"[case] two-tensor column-weighted row sums {method}", flush=True)
    O, st = gpu_contract(ctx, A, B, method=METHODS[method])
    h, v = O.hedges_array, O.val_array
    gs = np.bincount(h[:, 0], weights=v * column_weight(h[:, 1]), minlength=10000)
    nz = ws > 0
    assert np.array_equal(gs > 0, nz)
    assert max_rel(gs[nz], ws[nz]) <= 1e-10


def test_two_tensor_methods_agree_on_counts(ctx, c3):
    A, B = c3[0], c3[1]
    counts = []
    for m in (P.METHOD_TWO_PASS, P.METHOD_UPPER_BOUND):
        _, st = gpu_contract(ctx, A, B, method=m, download=False)
        counts.append(st.nnz_out)
    assert counts[0] == counts[1]


# ------------------------------------------------------------- FP32 path

F32_MODES = {"twosum": P.VAL_F32_TWOSUM, "f64acc": P.VAL_F32_F64ACC, "plain": P.VAL_F32_PLAIN}
ULP_REL = 2.0 ** -23


def with_f32_vals(T):
    return P.Tensor(T.hedges_array, T.val_array.astype(np.float32).astype(np.float64), T.dimensions_array, T.dim_tc,
                    T.num_key_dimensions)


def gpu_contract_f32(ctx, A, B, vt, **opts):
    dA, dB = P.DeviceTensor.upload_f32(A, ctx), P.DeviceTensor.upload_f32(B, ctx)
    st = P.Stats()
    O = P.tc_fksCuckoo(P.buildHashTable(dA, ctx), P.buildHashTable(dB, ctx), download=False, stats=st,
                       value_type=vt, **opts)
    h = O.download().hedges_array
    hi, lo = O.download_f32(with_lo=True)
    O.free()
    return h, hi, lo, st


def check_f32(mode, h, hi, exact, ref64, nterms):
    """hi against the exact result of the FP32 inputs (exact, sorted like
    h), and the error against the FP64 oracle (ref64) reported."""
    eh, ev = exact
    assert np.array_equal(h, eh)
    assert np.array_equal(h, ref64[0])
    rel = np.abs(hi.astype(np.float64) - ev) / np.abs(ev)
    if mode == "plain":
        # recursive summation of n positive products, each rounded once:
        # |hi - s| <= gamma_{n+1} s, gamma_k = k u / (1 - k u), u = 2^-24
        k = (nterms + 1).astype(np.float64) * 2.0 ** -24
        assert np.all(rel <= k / (1 - k) * 1.0001), float(np.max(rel / (k / (1 - k))))
    else:
        assert float(np.max(rel)) <= ULP_REL
    e64 = max_rel(hi, ref64[1])
    print(f"  {mode}: max rel err vs exact FP32-input result {float(np.max(rel)):.3e}, vs FP64 oracle {e64:.3e}")
    return e64


@pytest.fixture(scope="module")
def nips_f32(nips):
    A, B = nips[0], nips[1]
    oh, ov, _ = orc.tc_fksCuckoo(with_f32_vals(A), with_f32_vals(B))
    # terms of output entry (i, j) <= contracted keys rows i and j share <=
    # min(nonzeros of row i, of row j)
    rows = np.bincount(A.hedges_array[:, 0].astype(np.int64), minlength=2482)
    return sort_coo(oh, ov), rows


@pytest.mark.parametrize("mode", sorted(F32_MODES))
def test_nips_full_output_f32(ctx, nips, nips_f32, mode):
    A, B, ref64, _ = nips
    exact, rows = nips_f32
    print(f"[case] nips full size FP32 {mode}", flush=True)
    h, hi, lo, st = gpu_contract_f32(ctx, A, B, F32_MODES[mode])
    o = np.lexsort(h.T[::-1])
    h, hi = h[o], hi[o]
    nterms = np.minimum(rows[h[:, 0]], rows[h[:, 1]])
    check_f32(mode, h, hi, exact, ref64, nterms)


@pytest.fixture(scope="module")
def c3_f32(c3):
    A, B, _, rows, _ = c3
    sel = np.isin(A.hedges_array[:, 0], rows)
    Asub = P.Tensor(A.hedges_array[sel], A.val_array[sel], A.dimensions_array, A.dim_tc, A.num_key_dimensions)
    oh, ov, _ = orc.tc_fksCuckoo(with_f32_vals(Asub), with_f32_vals(B))
    rowcnt = np.bincount(A.hedges_array[:, 0].astype(np.int64), minlength=10000)
    return sort_coo(oh, ov), rowcnt


@pytest.mark.parametrize("mode", sorted(F32_MODES))
def test_two_tensor_f32(ctx, c3, c3_f32, mode):
    A, B, rs, rows, ref64 = c3
    exact, rowcnt = c3_f32
    print(f"[case] two-tensor full size FP32 {mode}", flush=True)
    h, hi, lo, st = gpu_contract_f32(ctx, A, B, F32_MODES[mode])
    # every row's (hi + lo) sum against the FP64 closed form on the original
    # values.  Rounding the inputs moves each positive product by at most
    # (1 + u)^2 - 1 < 2.0001 u (u = 2^-24); compensated sums add < 1e-9
    # relative, plain FP32 sums of n_i products at most gamma_{n_i + 1}
    # (n_i = the row's A nonzeros bounds the terms of each of its entries).
    u = 2.0 ** -24
    gs = np.bincount(h[:, 0], weights=hi.astype(np.float64) + lo.astype(np.float64), minlength=10000)
    nz = rs > 0
    assert np.array_equal(gs > 0, nz)
    k = (rowcnt[nz] + 1).astype(np.float64) * u
    bound = 2.0001 * u + (k / (1 - k) if mode == "plain" else 1e-9)
    rel = np.abs(gs[nz] - rs[nz]) / rs[nz]
    assert np.all(rel <= bound), float(np.max(rel / bound))
    sel = np.isin(h[:, 0], rows)
    hs, his = h[sel], hi[sel]
    o = np.lexsort(hs.T[::-1])
    hs, his = hs[o], his[o]
    check_f32(mode, hs, his, exact, ref64, rowcnt[hs[:, 0]])
