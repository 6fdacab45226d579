"""The batched workload at its full size: 256 independent self-contractions,
job i of shape gen_job_shape(1000 + i) (3 or 4 modes, dims log-uniform in
[64, 20000], nnz log-uniform in [1e3, 1e6], Zipf(1.1) coordinates, FP64
values), each contracted over every mode but mode 0 -- 36 M input nonzeros,
1.46e9 multiply-adds, 3.4e8 output nonzeros in all.

One batch plan (all key tables, the symbolic phase, the captured numeric
graph) runs once; then every input's values are doubled in place and the
graph is replayed.  Every job of both runs is compared with the oracle:
coordinates bit-exact, values within 1e-12 relative (the replay against four
times the oracle's values: doubling both operands scales every product, and
so every sum, by exactly four).

The oracle runs on 16 host threads (its calls release the interpreter lock,
and its hash-key random stream is per thread), and each thread also puts its
job's outputs in coordinate order (numpy sorts release the lock too).
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import partc_amd as P
from oracle import orc

pytestmark = pytest.mark.gpu
TOL = 1e-12
JOBS = 256
SEED0 = 1000


def job_tensor(i):
    d, dims, nnz = P.gen_job_shape(SEED0 + i)
    h, v = P.gen_coo(SEED0 + i, dims, nnz, P.DIST_ZIPF, 1.1)
    perm = np.array(list(range(1, d)) + [0], dtype=np.uint32)  # .meta: contracted modes first
    return P.Tensor(h, v, None, perm, d - 1)


def canon(h, v):
    """Output entries in coordinate order: (keys, values)."""
    assert h.shape[1] == 2
    key = (h[:, 0].astype(np.int64) << 32) | h[:, 1].astype(np.int64)
    o = np.argsort(key)
    return key[o], v[o]


def compare(gpu, ref, scale):
    """(ok, worst relative value error, message)."""
    gk, gv = canon(*gpu)
    ok, ov = canon(*ref)
    if gk.shape != ok.shape or not np.array_equal(gk, ok):
        return False, 0.0, f"coordinates differ ({gk.shape[0]} vs {ok.shape[0]} entries)"
    if gk.shape[0] > 1 and not np.all(gk[1:] != gk[:-1]):
        return False, 0.0, "repeated coordinates"
    ov = ov * scale
    rel = float(np.max(np.abs(gv - ov) / np.maximum(np.abs(ov), np.finfo(np.float64).tiny))) if gv.size else 0.0
    return rel <= TOL, rel, f"max rel {rel:.3e}"


def test_full_batch_and_replay_match_oracle(ctx):
    print(f"[case] batch of {JOBS} uneven self-contractions at full size", flush=True)
    Ts = [job_tensor(i) for i in range(JOBS)]
    dev = [P.DeviceTensor.upload(T, ctx) for T in Ts]
    plan = P.BatchPlan(dev, dev, ctx)
    st = plan.run()
    print(f"  flops {st.flops} nnz_out {st.nnz_out} bins {st.bins} items {st.queue_items}", flush=True)
    assert st.jobs == JOBS
    first = []
    for j in range(JOBS):
        O = plan.output(j)
        first.append((O.hedges_array, O.val_array))
    for j, T in enumerate(Ts):
        dev[j].set_values(2.0 * T.val_array)
    st = plan.run()
    assert st.runs == 2
    replay = []
    for j in range(JOBS):
        O = plan.output(j)
        replay.append((O.hedges_array, O.val_array))
    plan.free()
    print("  both runs downloaded; oracle on 16 threads", flush=True)
    assert sum(h.shape[0] for h, _ in first) == st.nnz_out

    def check(j):
        oh, ov, info = orc.contract_meta(Ts[j], Ts[j])
        r1 = compare(first[j], (oh, ov), 1.0)
        r2 = compare(replay[j], (oh, ov), 4.0)
        first[j] = replay[j] = None
        return j, info["flops"], r1, r2

    flops = 0
    worst = 0.0
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        for n, (j, f, r1, r2) in enumerate(ex.map(check, range(JOBS))):
            flops += f
            assert r1[0], f"job {j}: {r1[2]}"
            assert r2[0], f"replay job {j}: {r2[2]}"
            worst = max(worst, r1[1], r2[1])
            if n % 32 == 31:
                print(f"  {n + 1} jobs checked", flush=True)
    assert flops == st.flops
    print(f"  all {JOBS} jobs match, both runs; worst rel err {worst:.2e}", flush=True)
