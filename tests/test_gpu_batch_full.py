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
and its hash-key random stream is per thread); the outputs, 3.4e8 entries a
run, are put in coordinate order by sorting on the GPU with torch, which is
test plumbing only.
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
    """Output entries in coordinate order, on the GPU: (keys, values)."""
    import torch

    assert h.shape[1] == 2
    key = (h[:, 0].astype(np.int64) << 32) | h[:, 1].astype(np.int64)
    k = torch.from_numpy(key).to("cuda:0")
    sk, idx = torch.sort(k)
    return sk, torch.from_numpy(np.ascontiguousarray(v)).to("cuda:0")[idx]


def compare(gpu, ref, scale, label):
    import torch

    gk, gv = canon(*gpu)
    ok, ov = canon(*ref)
    assert gk.shape == ok.shape, (label, tuple(gk.shape), tuple(ok.shape))
    assert torch.equal(gk, ok), label
    if gk.numel() > 1:
        assert bool((gk[1:] != gk[:-1]).all()), label  # unique coordinates
    ov = ov * scale
    rel = ((gv - ov).abs() / ov.abs().clamp_min(np.finfo(np.float64).tiny)).max().item() if gv.numel() else 0.0
    assert rel <= TOL, (label, rel)
    return rel


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

    def ref(j):
        oh, ov, info = orc.contract_meta(Ts[j], Ts[j])
        return j, oh, ov, info["flops"]

    flops = 0
    worst = 0.0
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        for n, (j, oh, ov, f) in enumerate(ex.map(ref, range(JOBS))):
            flops += f
            worst = max(worst, compare(first[j], (oh, ov), 1.0, f"job {j}"))
            worst = max(worst, compare(replay[j], (oh, ov), 4.0, f"replay job {j}"))
            first[j] = replay[j] = None
            if n % 32 == 31:
                print(f"  {n + 1} jobs checked", flush=True)
    assert flops == st.flops
    print(f"  all {JOBS} jobs match, both runs; worst rel err {worst:.2e}", flush=True)
