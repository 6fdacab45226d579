"""Batched uneven contractions (ptc_batch_plan / ptc_batch_run /
ptc_contract_batched): every job's output against the oracle (coordinates
bit-exact after sorting, values within 1e-12 relative), against the
single-contraction entry point, and across replays of the captured numeric
graph after the input values change in place."""
import numpy as np
import pytest

import partc_amd as P
from oracle import orc
from tests.helpers import max_rel, sort_coo

pytestmark = pytest.mark.gpu
TOL = 1e-12


def meta_self(h, v, contracted):
    """A .meta-style tensor (contracted modes first, num_key_dimensions = c)
    for a self-contraction over `contracted`: both operands of
    ptc_contract."""
    d = h.shape[1]
    free = [m for m in range(d) if m not in contracted]
    perm = np.array(list(contracted) + free, dtype=np.uint32)
    return P.Tensor(h, v, None, perm, len(contracted))


def job_inputs():
    """Uneven jobs: shapes of the batched workload family (nnz capped so the
    oracle stays fast), plus a dense job with rows split over many
    workgroups and jobs with column spaces above the dense limit (their
    large rows run as dense column windows -- three for 40000 columns --
    and their small rows in LDS hash tables)."""
    jobs = []
    for i in range(10):
        d, dims, nnz = P.gen_job_shape(1000 + i)
        nnz = min(nnz, 20000)
        h, v = P.gen_coo(1000 + i, dims, nnz, P.DIST_ZIPF)
        jobs.append((h, v, list(range(1, d))))
    h, v = P.gen_coo(41, [60, 50, 40], 30000, P.DIST_ZIPF)  # heavy rows
    jobs.append((h, v, [1, 2]))
    h, v = P.gen_coo(42, [12000, 30, 40], 20000, P.DIST_ZIPF)  # 12000 columns: hash accumulators
    jobs.append((h, v, [1, 2]))
    h, v = P.gen_coo(43, [9000, 20, 30, 7], 15000, P.DIST_ZIPF)
    jobs.append((h, v, [1, 2, 3]))
    h, v = P.gen_coo(46, [40000, 30, 40], 40000, P.DIST_ZIPF)  # 40000 columns: three dense column windows
    jobs.append((h, v, [1, 2]))
    return jobs


def oracle_of(T):
    oh, ov, _ = orc.tc_fksCuckoo(P.make_free_first(T), T)
    return sort_coo(oh, ov)


def check(O, ref, label):
    h, v = sort_coo(O.hedges_array, O.val_array)
    assert h.shape == ref[0].shape, (label, h.shape, ref[0].shape)
    assert np.array_equal(h, ref[0]), label
    assert max_rel(v, ref[1]) <= TOL, label


@pytest.fixture(scope="module")
def jobs():
    return [meta_self(h, v, c) for h, v, c in job_inputs()]


@pytest.mark.parametrize("chunk", [0, 512])
def test_batch_matches_oracle(ctx, jobs, chunk):
    print(f"[case] batch of {len(jobs)} chunk={chunk}", flush=True)
    st = P.BatchStats()
    outs = P.contract_batched(jobs, jobs, ctx, stats=st, chunk_flops=chunk)
    print(f"  bins={st.bins} items={st.queue_items} nodes={st.graph_nodes} flops={st.flops}")
    assert st.jobs == len(jobs) and st.bins >= 2
    for j, (T, O) in enumerate(zip(jobs, outs)):
        check(O, oracle_of(T), f"job {j}")


def test_batch_equals_single_contractions(ctx, jobs):
    print("[case] batch vs contract()", flush=True)
    outs = P.contract_batched(jobs, jobs, ctx)
    for j, (T, O) in enumerate(zip(jobs, outs)):
        S = P.contract(T, T, ctx)
        h1, v1 = sort_coo(O.hedges_array, O.val_array)
        h2, v2 = sort_coo(S.hedges_array, S.val_array)
        assert np.array_equal(h1, h2), j
        assert max_rel(v1, v2) <= TOL, j


def test_batch_graph_replays_new_values(ctx, jobs):
    """The captured numeric graph, launched again after every tensor's values
    changed in place, gives the contraction of the new values -- checked on
    each of several replays with host work in between."""
    print("[case] batch replay", flush=True)
    dev = [P.DeviceTensor.upload(T, ctx) for T in jobs]
    plan = P.BatchPlan(dev, dev, ctx)
    plan.run()
    for j, T in enumerate(jobs):
        check(plan.output(j), oracle_of(T), f"first run job {j}")
    rng = np.random.default_rng(7)
    for rep in range(3):
        news = []
        for j, T in enumerate(jobs):
            nv = rng.uniform(0.5, 2.0, T.N)
            dev[j].set_values(nv)
            news.append(P.Tensor(T.hedges_array, nv, T.dimensions_array, T.dim_tc, T.num_key_dimensions))
        st = plan.run()
        assert st.runs == rep + 2
        for j, T2 in enumerate(news):
            check(plan.output(j), oracle_of(T2), f"replay {rep} job {j}")
    plan.free()


def test_autotuned_batch_matches_oracle(ctx, jobs):
    """The tuner's chosen thresholds change how work is binned and
    accumulated, never the result."""
    print("[case] batch autotune", flush=True)
    tuned, rep = P.autotune_batch(jobs, jobs, ctx)
    print(f"  tuned {tuned} report {rep}")
    assert rep["trials"] >= 5 and rep["best_ms"] <= rep["default_ms"]
    outs = P.contract_batched(jobs, jobs, ctx, **tuned)
    for j, (T, O) in enumerate(zip(jobs, outs)):
        check(O, oracle_of(T), f"tuned job {j}")


def test_batch_errors(ctx, jobs):
    with pytest.raises(P.PtcError) as ei:
        P.contract_batched(jobs[:2], jobs[:2], ctx, value_type=P.VAL_F32_TWOSUM)
    assert ei.value.code == -6
    h, v = P.gen_coo(5, [20, 30, 40], 500, P.DIST_UNIFORM)
    A = meta_self(h, v, [1, 2])
    B = meta_self(h, v, [2])
    with pytest.raises(P.PtcError) as ei:
        P.contract_batched([A], [B], ctx)
    assert ei.value.code == -1


def test_batch_wide_dense_accumulators(ctx):
    """A dense job whose accumulator and flag bytes fill most of a CU's LDS
    (15000 columns with the dense limit raised to 16384): the batch keeps
    flags for every job, so its workgroup size has to be chosen with them
    counted."""
    print("[case] batch wide dense", flush=True)
    h, v = P.gen_coo(44, [15000, 30, 40], 20000, P.DIST_ZIPF)
    wide = meta_self(h, v, [1, 2])
    h, v = P.gen_coo(45, [50, 40, 30], 8000, P.DIST_ZIPF)
    small = meta_self(h, v, [1, 2])
    batch = [wide, small]
    st = P.BatchStats()
    outs = P.contract_batched(batch, batch, ctx, stats=st, dense_max_cols=16384)
    for j, (T, O) in enumerate(zip(batch, outs)):
        check(O, oracle_of(T), f"job {j}")
