"""The learned single-pass method: the row-size network's GPU kernels
(gradients against the CPU reference oracle/mlp_ref.py), its training, and
the contraction it sizes (parity with the oracle, exact fallback)."""
import numpy as np
import pytest

import partc_amd as P
from oracle import mlp_ref, orc
from tests.helpers import golden_cases, max_rel, sort_coo, tc_inputs

pytestmark = pytest.mark.gpu
TOL = 1e-12


def test_gradients_match_cpu_reference(ctx):
    rng = np.random.default_rng(5)
    m = P.RowSizeModel(ctx, seed=3)
    w = m.params().astype(np.float64)
    n = 300
    X = rng.uniform(0, 2, (n, P.MLP_IN)).astype(np.float32)
    ub = rng.uniform(5, 5000, n).astype(np.float32)
    y = (ub * rng.uniform(0.05, 1.0, n)).astype(np.float32)
    g, loss = m.grad(X, ub, y)
    rl, rg = mlp_ref.loss_and_grad(w, X.astype(np.float64), ub.astype(np.float64), y.astype(np.float64))
    assert abs(loss - rl) <= 1e-4 * max(1.0, abs(rl))
    scale = np.max(np.abs(rg))
    assert np.max(np.abs(g - rg)) <= 2e-4 * scale, np.max(np.abs(g - rg)) / scale
    # predictions
    p = m.predict(X, ub)
    assert max_rel(p, mlp_ref.predict(w, X, ub)) <= 1e-4


def _profiles(ctx, seeds):
    Xs, us, ys = [], [], []
    for s in seeds:
        h, v = P.gen_coo(s, [300, 60, 80, 9], 40000, P.DIST_ZIPF)
        A, B = P.self_contraction_inputs(h, v, [1, 2, 3])
        X, ub, nnz = P.row_profile(P.buildHashTable(A, ctx), P.buildHashTable(B, ctx))
        keep = ub >= 1
        Xs.append(X[keep]); us.append(ub[keep]); ys.append(nnz[keep].astype(np.float32))
        h, v = P.gen_coo(s + 1, [80, 3000, 4000], 30000, P.DIST_ZIPF)
        A, B = P.self_contraction_inputs(h, v, [0])
        X, ub, nnz = P.row_profile(P.buildHashTable(A, ctx), P.buildHashTable(B, ctx))
        keep = ub >= 1
        Xs.append(X[keep]); us.append(ub[keep]); ys.append(nnz[keep].astype(np.float32))
    return np.concatenate(Xs), np.concatenate(us), np.concatenate(ys)


@pytest.fixture(scope="module")
def trained(ctx):
    X, ub, y = _profiles(ctx, [900, 902, 904, 906])
    m = P.RowSizeModel(ctx, seed=11)
    losses = m.train(X, ub, y, epochs=60, batch=1024, lr=3e-3)
    return m, losses, (X, ub, y)


def test_profile_counts_are_exact(ctx):
    h, v = P.gen_coo(31, [100, 40, 50], 20000, P.DIST_ZIPF)
    A, B = P.self_contraction_inputs(h, v, [1, 2])
    HA, HB = P.buildHashTable(A, ctx), P.buildHashTable(B, ctx)
    X, ub, nnz = P.row_profile(HA, HB)
    oh, ov, _ = orc.tc_fksCuckoo(A, B)
    exp = np.bincount(oh[:, 0], minlength=100)
    # rows are in table order: compare the multisets of per-row counts
    assert sorted(nnz.tolist()) == sorted(exp[exp > 0].tolist() + [0] * (nnz.shape[0] - int((exp > 0).sum())))
    assert np.all(nnz <= ub)


def test_training_reduces_loss(trained):
    m, losses, _ = trained
    assert losses[-1] < 0.5 * losses[0], losses


def test_trained_network_beats_untrained(ctx, trained):
    """What the network gains: at the same capacity margin, the trained
    network's capacities cover far more rows (fewer exact fallbacks) than an
    untrained one, with total capacity well under the flop upper bound."""
    m, _, _ = trained
    X, ub, y = _profiles(ctx, [950])
    fresh = P.RowSizeModel(ctx, seed=12)

    def stats(model):
        cap = np.minimum(np.ceil(model.predict(X, ub) * 1.15), ub)
        return float(np.mean(cap < y)), float(cap.sum() / y.sum())

    fb_t, over_t = stats(m)
    fb_u, over_u = stats(fresh)
    print(f"trained fallback {fb_t:.3f} cap/exact {over_t:.2f}; untrained {fb_u:.3f} {over_u:.2f}; "
          f"upper bound/exact {ub.sum() / y.sum():.2f}")
    assert fb_t < 0.5 * fb_u or fb_t < 0.02
    assert over_t < ub.sum() / y.sum()


@pytest.mark.parametrize("name", sorted(golden_cases()))
def test_learned_golden(ctx, trained, name):
    m, _, _ = trained
    case = golden_cases()[name]
    A, B = tc_inputs(case)
    st = P.Stats()
    O = P.tc_fksCuckoo(P.buildHashTable(A, ctx), P.buildHashTable(B, ctx), method=P.METHOD_LEARNED, mlp=m, stats=st)
    h, v = sort_coo(O.hedges_array, O.val_array)
    assert np.array_equal(h, case["coords"])
    assert max_rel(v, case["vals_ordered"]) <= TOL


@pytest.mark.parametrize("scale", [0.01, 0.5, 1.15, 4.0])
@pytest.mark.parametrize("shape", ["dense", "hash"])
def test_learned_vs_oracle_with_fallback(ctx, trained, scale, shape):
    """Tiny margins force most rows through the exact fallback; results stay exact."""
    m, _, _ = trained
    if shape == "dense":
        h, v = P.gen_coo(960, [300, 60, 80, 9], 40000, P.DIST_ZIPF)
        A, B = P.self_contraction_inputs(h, v, [1, 2, 3])
    else:
        h, v = P.gen_coo(961, [80, 3000, 4000], 30000, P.DIST_ZIPF)
        A, B = P.self_contraction_inputs(h, v, [0])
    print(f"[case] learned {shape} scale={scale}", flush=True)
    st = P.Stats()
    O = P.tc_fksCuckoo(P.buildHashTable(A, ctx), P.buildHashTable(B, ctx), method=P.METHOD_LEARNED, mlp=m,
                       predict_scale=scale, stats=st)
    h1, v1 = sort_coo(O.hedges_array, O.val_array)
    oh, ov, _ = orc.tc_fksCuckoo(A, B)
    oh, ov = sort_coo(oh, ov)
    assert np.array_equal(h1, oh)
    assert max_rel(v1, ov) <= TOL
    if scale <= 0.01:
        assert st.fallback_rows > 0


@pytest.mark.parametrize("shape", [([6000, 40, 30, 20], 60000, [1, 2, 3]), ([12000, 30, 40], 20000, [1, 2])])
def test_learned_allocates_less_than_upper_bound(ctx, shape):
    """What the network gains over the method without it (PTC_METHOD_UPPER_BOUND,
    capacity min(flops, columns) per row): on a contraction whose rows hold
    far fewer distinct columns than multiply-adds, the learned capacities
    size the output arrays to about the true count instead of the flop
    bound, with the same result (the built-in model, trained on other
    tensors).  Both shapes have a rows x columns bound above the size below
    which a single pass over dense rows runs ahead of its preprocessing
    totals with the output sized by that bound (512 MiB), so the capacities
    decide the allocation.  On dense outputs (the nips-shape and two-tensor
    workloads) the two capacities coincide, and the learned method only adds
    its features and inference."""
    if not P.learned_available():
        pytest.skip("no trained row-size model in partc_amd/models")
    dims, nnz, contracted = shape
    h, v = P.gen_coo(42, dims, nnz, P.DIST_ZIPF)
    A, B = P.self_contraction_inputs(h, v, contracted)
    print(f"[case] learned vs upper bound capacity {dims} {nnz}", flush=True)
    res = {}
    for name, m in (("upper_bound", P.METHOD_UPPER_BOUND), ("learned", P.METHOD_LEARNED)):
        st = P.Stats()
        O = P.tc_fksCuckoo(P.buildHashTable(A, ctx), P.buildHashTable(B, ctx), stats=st, method=m)
        res[name] = (sort_coo(O.hedges_array, O.val_array), st.out_capacity, st.nnz_out, st.fallback_rows, st.rows)
    (hu, vu), cap_u, n_u, _, _ = res["upper_bound"]
    (hl, vl), cap_l, n_l, fb, rows = res["learned"]
    oh, ov, _ = orc.tc_fksCuckoo(A, B)
    oh, ov = sort_coo(oh, ov)
    assert np.array_equal(hu, oh) and np.array_equal(hl, oh)
    assert max_rel(vu, ov) <= TOL and max_rel(vl, ov) <= TOL
    print(f"  nnz {n_l}: capacity learned {cap_l} ({cap_l / n_l:.2f}x), upper bound {cap_u} ({cap_u / n_u:.2f}x); "
          f"fallback rows {fb} of {rows}")
    assert cap_l <= 1.3 * n_l          # about the true count ...
    assert cap_l <= 0.85 * cap_u       # ... and well below the flop bound
    assert fb <= 0.05 * rows
