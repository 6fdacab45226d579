"""Benchmark of the hot path: the nips-shape 4-mode self-contraction.

Workload (one GPU): a seeded 4-mode COO tensor, dims 2482 x 2862 x 14036 x 17,
3,101,609 unique nonzeros with coordinates drawn per mode from Zipf(1.1) and
values U(0,1] in FP64, contracted with itself over modes {1,2,3} into a sparse
2482 x 2482 output.

A step is one whole contraction from COO tensors resident in HBM to the COO
output in HBM: the two key tables (B on its contracted modes, A on its free
modes) and the contraction itself -- what main.cpp:115-118 times as HT_B
insertion + HT_A insertion + tc_fksCuckoo.  The value is useful multiply-adds
(F = sum_k |A_k| |B_k|) per second of step time.

Printed: one JSON line (rank 0) with the contract's fields plus
  methods      each method's step time and MAD/s (two-pass exact, upper-bound
               single pass, learned single pass when a model is available);
  roofline     the dominant kernel (the numeric row pass) against HBM peak:
               algorithmic bytes per launch / HIP-event kernel duration;
  cpu_baseline the oracle restatement of the reference (one thread, as the
               reference's numeric loop runs) on the same tensor.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
DIMS = [2482, 2862, 14036, 17]
NNZ = 3101609
CONTRACTED = [1, 2, 3]
SEED = 2
WORKLOAD = ("nips-shape 4-mode self-contraction: dims 2482x2862x14036x17, 3101609 unique nnz, "
            "Zipf(1.1) per mode, FP64 U(0,1], contract modes {1,2,3} -> 2482x2482")


def algorithmic_bytes(nnz_a, d_a, c, d_b, F, nnz_c, d_o):
    """Slice-streaming model of the contraction: read A once, one key probe
    per A nonzero, stream one B-slice entry (free index + value) per
    multiply-add, write C once."""
    return (nnz_a * (4 * d_a + 8) + nnz_a * (4 * c + 8) + F * (4 * (d_b - c) + 8) + nnz_c * (4 * d_o + 8))


def kernel_bytes(npos, F, d_b, c, nnz_c, d_o):
    """Algorithmic bytes of one numeric row pass: per A position its flop
    prefix (8), slice offset (4) and value (8); per multiply-add one B entry
    (column key 4*(d_B-c) + value 8); per output entry its coordinates and
    value."""
    return npos * 20 + F * (4 * (d_b - c) + 8) + nnz_c * (4 * d_o + 8)


def streamed_bytes(npos, F, nnz_c, d_o):
    """Bytes the dense numeric pass actually has to stream: the same terms as
    kernel_bytes with the 16-bit column keys dense rows read (their column
    space is below 2^16), so 10 bytes per multiply-add instead of 12."""
    return npos * 20 + F * (2 + 8) + nnz_c * (4 * d_o + 8)


def run_method(P, ctx, dA, dB, method, steps, warmup, value_type=0, stat_steps=3):
    """Times `steps` whole contractions with phase timing off (as a caller
    runs them: no HIP events between phases); then `stat_steps` more with
    phase timing on give the per-phase statistics and the numeric kernel's
    event-bracketed duration.  Returns per-step ms and stats."""
    import ctypes as C

    L = P.lib()
    opts = P.options(method=method, value_type=value_type)
    st = P.Stats()

    def one():
        hb, ha, out = C.c_void_p(), C.c_void_p(), C.c_void_p()
        rc = L.ptc_hashtable_build(ctx.handle, dB._h, C.byref(hb))
        if rc == 0:
            rc = L.ptc_hashtable_build(ctx.handle, dA._h, C.byref(ha))
        if rc == 0:
            rc = L.ptc_tc_fksCuckoo(ctx.handle, ha, hb, dA._h, dB._h, C.byref(opts), C.byref(out), C.byref(st))
        L.ptc_dtensor_free(out)
        L.ptc_hashtable_free(ha)
        L.ptc_hashtable_free(hb)
        if rc != 0:
            raise P.PtcError(rc, ctx.last_error())

    ctx.set_timing(False)
    for _ in range(warmup):
        one()
    ctx.synchronize()
    t0 = time.perf_counter()
    per = []
    for _ in range(steps):
        s0 = time.perf_counter()
        one()
        per.append(time.perf_counter() - s0)
    ctx.synchronize()
    total = time.perf_counter() - t0
    # phase statistics and kernel durations, outside the timed steps
    ctx.set_timing(True)
    kern_ms, kern_launch = 0.0, 0
    for _ in range(max(1, stat_steps)):
        one()
        kern_ms += st.ms_kernel_numeric
        kern_launch += st.kernel_launches_numeric
    ctx.set_timing(False)
    return {
        "ms_per_step": 1e3 * total / steps,
        "ms_min": 1e3 * min(per),
        "stats": st.as_dict(),
        "kernel_ms_per_launch": kern_ms / max(kern_launch, 1),
        "kernel_launches": kern_launch,
    }


def f32_errors(P, ctx, dA, dB, method, value_type, ref):
    """One untimed FP32 contraction; max relative error of its values (hi =
    the sum rounded to FP32) against the FP64 oracle's output `ref`."""
    HB, HA = P.buildHashTable(dB, ctx), P.buildHashTable(dA, ctx)
    O = P.tc_fksCuckoo(HA, HB, download=False, method=method, value_type=value_type)
    h = O.download().hedges_array
    hi = O.download_f32().astype(np.float64)
    o = np.lexsort(h.T[::-1])
    h, hi = h[o], hi[o]
    rh, rv = ref
    if not np.array_equal(h, rh):
        return {"coords_equal": False}
    rel = np.abs(hi - rv) / np.abs(rv)
    return {"coords_equal": True, "max_rel_err": float(rel.max()), "mean_rel_err": float(rel.mean())}


def host_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_threads_variant(P, A, B, threads):
    """The oracle on `threads` host threads at once, A's output rows dealt
    round-robin by row index: each thread builds its own HT_B and contracts
    its rows (the optional fixed-threads variant; the reference itself runs
    every phase on one thread)."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import orc

    rows = A.hedges_array[:, A.dim_tc[0]]
    parts = []
    for t in range(threads):
        sel = rows % threads == t
        parts.append(P.Tensor(A.hedges_array[sel], A.val_array[sel], A.dimensions_array, A.dim_tc,
                              A.num_key_dimensions))
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        res = list(ex.map(lambda a: orc.tc_fksCuckoo(a, B), parts))
    wall = time.perf_counter() - t0
    F = sum(r[2]["flops"] for r in res)
    return {"threads": threads, "value": F / wall, "unit": "MAD/s", "seconds": wall,
            "nnz_out": int(sum(r[0].shape[0] for r in res)),
            "note": "rows dealt to threads by row index mod threads; HT_B built once per thread"}


def cpu_baseline(P, h, v):
    """The oracle restatement of the reference on this host: the whole
    contraction, one thread -- the reference runs every phase on one thread
    whatever OMP_NUM_THREADS says (its only parallel region, the numeric
    loop, takes its team size from omp_get_num_threads() called outside a
    parallel region, which is 1: fksCuckooTC.hpp:171-173), so runs at
    OMP_NUM_THREADS=4 or =nproc are this run.  Beside it: the fixed-threads
    variant on the host's cores (at most 16), and the CPU configuration of
    the reference's own scale (3-mode 200x300x400, 50k uniform nnz,
    contract {1,2})."""
    from oracle import orc

    A, B = P.self_contraction_inputs(h, v, CONTRACTED)
    t0 = time.perf_counter()
    oh, ov, info = orc.tc_fksCuckoo(A, B)
    wall = time.perf_counter() - t0
    F = info["flops"]
    ncores = host_cores()
    out = {
        "value": F / wall,
        "unit": "MAD/s",
        "cores": 1,
        "host_cores": ncores,
        "kind": "port",
        "sample": (f"whole workload (all {info['rows']} rows), single thread: HT_B {info['t_build_b']:.2f}s, "
                   f"HT_A {info['t_build_a']:.2f}s, preprocessing {info['t_preprocess']:.2f}s, "
                   f"contraction {info['t_contract']:.2f}s; nnz_out {oh.shape[0]}"),
        "nnz_out": int(oh.shape[0]),
        "seconds": wall,
        "omp": "OMP_NUM_THREADS=4 and =nproc give this same single-thread run (fksCuckooTC.hpp:171-173)",
    }
    out["threads_variant"] = cpu_threads_variant(P, A, B, max(1, min(ncores, 16)))
    h1, v1 = P.gen_coo(1, [200, 300, 400], 50000, P.DIST_UNIFORM)
    A1, B1 = P.self_contraction_inputs(h1, v1, [1, 2])
    t0 = time.perf_counter()
    o1h, _, i1 = orc.tc_fksCuckoo(A1, B1)
    w1 = time.perf_counter() - t0
    out["c1"] = {"workload": "3-mode 200x300x400, 50k uniform nnz, seed 1, contract {1,2}", "seconds": w1,
                 "MADps": i1["flops"] / w1, "flops": int(i1["flops"]), "nnz_out": int(o1h.shape[0]),
                 "phases_s": {k: round(i1[k], 5) for k in ("t_build_b", "t_build_a", "t_preprocess", "t_contract")}}
    return out, _sorted(oh, ov)


def gpu_c1(P, ctx, steps=20):
    """The CPU configuration on the GPU (launch-bound at this size)."""
    h, v = P.gen_coo(1, [200, 300, 400], 50000, P.DIST_UNIFORM)
    A, B = P.self_contraction_inputs(h, v, [1, 2])
    dA, dB = P.DeviceTensor.upload(A, ctx), P.DeviceTensor.upload(B, ctx)
    r = run_method(P, ctx, dA, dB, P.METHOD_TWO_PASS, steps, 3)
    return {"ms_per_step": round(r["ms_per_step"], 4), "MADps": r["stats"]["flops"] / (r["ms_per_step"] * 1e-3)}


def _sorted(h, v):
    o = np.lexsort(h.T[::-1])
    return h[o], v[o]


def copy_calibration(mb: int, reps: int = 5) -> None:
    """A plain device-to-device copy of `mb` MiB (hipMemcpy), for scaling
    the FETCH_SIZE / WRITE_SIZE counters of rocprofv3 --pmc runs."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mb << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n, "reps": reps}))


BATCH_JOBS = 256
BATCH_SEED0 = 1000
BATCH_WORKLOAD = ("256 uneven self-contractions: job i seeded 1000+i, 3 or 4 modes, dims log-uniform in "
                  "[64, 20000], nnz log-uniform in [1e3, 1e6], Zipf(1.1), FP64 U(0,1], every mode but mode 0 "
                  "contracted")


def batch_bench(P, args):
    """Config of many uneven contractions: the 256 jobs one ptc_contract call
    after another, against one batch (ptc_contract_batched: all key tables,
    batch-wide read-backs, persistent row passes, captured numeric graph) and
    against a replay of the batch's numeric graph."""
    import ctypes as C

    L = P.lib()
    ctx = P.Context(0)
    t0 = time.perf_counter()
    dev = []
    for i in range(BATCH_JOBS):
        d, dims, nnz = P.gen_job_shape(BATCH_SEED0 + i)
        h, v = P.gen_coo(BATCH_SEED0 + i, dims, nnz, P.DIST_ZIPF, 1.1)
        perm = np.array(list(range(1, d)) + [0], dtype=np.uint32)
        dev.append(P.DeviceTensor.upload(P.Tensor(h, v, None, perm, d - 1), ctx))
    print(f"# batch inputs ready in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    opts = P.options()
    arr = (C.c_void_p * BATCH_JOBS)(*[x._h.value for x in dev])

    def sequential(o=None):
        o = opts if o is None else o
        F = 0
        for x in dev:
            out, st = C.c_void_p(), P.Stats()
            rc = L.ptc_contract(ctx.handle, x._h, x._h, C.byref(o), C.byref(out), C.byref(st))
            if rc:
                raise P.PtcError(rc, ctx.last_error())
            F += st.flops
            L.ptc_dtensor_free(out)
        ctx.synchronize()
        return F

    def batched():
        outs = (C.c_void_p * BATCH_JOBS)()
        st = P.BatchStats()
        rc = L.ptc_contract_batched(ctx.handle, BATCH_JOBS, arr, arr, C.byref(opts), outs, C.byref(st))
        if rc:
            raise P.PtcError(rc, ctx.last_error())
        for o in outs:
            L.ptc_dtensor_free(C.c_void_p(o))
        ctx.synchronize()
        return st

    res = {}
    for name, fn in (("sequential", sequential), ("batched", batched)):
        for _ in range(args.warmup):
            fn()
        ts = []
        for _ in range(args.steps):
            a = time.perf_counter()
            r = fn()
            ts.append(time.perf_counter() - a)
        res[name] = (1e3 * float(np.median(ts)), r)
        print(f"# {name}: {res[name][0]:.2f} ms", file=sys.stderr, flush=True)
    F = res["sequential"][1]
    st = res["batched"][1]
    # numeric replays of one plan
    plan = P.BatchPlan(dev, dev, ctx)
    for _ in range(args.warmup):
        plan.run()
    reps = []
    for _ in range(args.steps):
        a = time.perf_counter()
        plan.run()
        reps.append(time.perf_counter() - a)
    rs = plan.stats
    replay_ms = 1e3 * float(np.median(reps))
    plan.free()
    # the autotuner's thresholds for this batch, then the batch again under them
    tuned, trep = P.autotune_batch(dev, dev, ctx)
    topts = P.options(**tuned)
    print(f"# autotune: {trep}", file=sys.stderr, flush=True)

    def batched_tuned():
        outs = (C.c_void_p * BATCH_JOBS)()
        st = P.BatchStats()
        rc = L.ptc_contract_batched(ctx.handle, BATCH_JOBS, arr, arr, C.byref(topts), outs, C.byref(st))
        if rc:
            raise P.PtcError(rc, ctx.last_error())
        for o in outs:
            L.ptc_dtensor_free(C.c_void_p(o))
        ctx.synchronize()
        return st

    tts = []
    for _ in range(args.warmup):
        batched_tuned()
    for _ in range(args.steps):
        a = time.perf_counter()
        tst = batched_tuned()
        tts.append(time.perf_counter() - a)
    tuned_ms = 1e3 * float(np.median(tts))
    print(f"# batched (tuned): {tuned_ms:.2f} ms", file=sys.stderr, flush=True)
    # the sequential loop under the same tuned thresholds (a fair comparison)
    sts = []
    for _ in range(args.steps):
        a = time.perf_counter()
        sequential(topts)
        sts.append(time.perf_counter() - a)
    seq_tuned_ms = 1e3 * float(np.median(sts))
    print(f"# sequential (tuned): {seq_tuned_ms:.2f} ms", file=sys.stderr, flush=True)
    seq_ms, bat_ms = res["sequential"][0], res["batched"][0]
    out = {
        "metric": "useful multiply-adds/s, 256 uneven self-contractions (one batch vs 256 sequential calls)",
        "value": F / (min(bat_ms, tuned_ms) * 1e-3),
        "unit": "MAD/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": min(bat_ms, tuned_ms),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator; no dataset)",
        "config": {"workload": BATCH_WORKLOAD, "flops": int(F), "nnz_out": int(st.nnz_out),
                   "step": "all 256 contractions from device-resident COO inputs to device COO outputs"},
        "sequential": {"ms": round(seq_ms, 3), "MADps": F / (seq_ms * 1e-3),
                       "ms_with_tuned_options": round(seq_tuned_ms, 3)},
        "batched": {"ms": round(bat_ms, 3), "MADps": F / (bat_ms * 1e-3), "speedup_vs_sequential": seq_ms / bat_ms,
                    "bins": int(st.bins), "queue_items": int(st.queue_items), "graph_nodes": int(st.graph_nodes),
                    "ms_build_device": round(st.ms_build, 3), "ms_symbolic_device": round(st.ms_symbolic, 3),
                    "ms_numeric_device": round(st.ms_numeric, 3)},
        "numeric_replay": {"ms": round(replay_ms, 3), "ms_device": round(rs.ms_numeric, 3),
                           "MADps": F / (replay_ms * 1e-3)},
        "autotuned": {"ms": round(tuned_ms, 3), "MADps": F / (tuned_ms * 1e-3),
                      "speedup_vs_sequential": seq_ms / tuned_ms,
                      "speedup_vs_sequential_tuned": seq_tuned_ms / tuned_ms,
                      "ms_numeric_device": round(tst.ms_numeric, 3),
                      "ms_symbolic_device": round(tst.ms_symbolic, 3), "bins": int(tst.bins), "tuner": trep},
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def batch_plan_bench(P, args):
    """One plan of the 256-job batch under fixed thresholds (--dense-max,
    --lds-hash, --chunk; 0 = defaults) and `steps` launches of its numeric
    graph: a short run for kernel traces and counters of the batch's row
    passes."""
    ctx = P.Context(0)
    dev = []
    for i in range(BATCH_JOBS):
        d, dims, nnz = P.gen_job_shape(BATCH_SEED0 + i)
        h, v = P.gen_coo(BATCH_SEED0 + i, dims, nnz, P.DIST_ZIPF, 1.1)
        perm = np.array(list(range(1, d)) + [0], dtype=np.uint32)
        dev.append(P.DeviceTensor.upload(P.Tensor(h, v, None, perm, d - 1), ctx))
    print("# batch inputs ready", file=sys.stderr, flush=True)
    plan = P.BatchPlan(dev, dev, ctx, dense_max_cols=args.dense_max, lds_hash_max=args.lds_hash,
                       chunk_flops=args.chunk)
    st = plan.stats
    runs = []
    for _ in range(max(args.steps, 1)):
        plan.run()
        runs.append(plan.stats.ms_numeric)
    out = {"config": {"workload": BATCH_WORKLOAD, "dense_max_cols": args.dense_max, "lds_hash_max": args.lds_hash,
                      "chunk_flops": args.chunk},
           "flops": int(st.flops), "nnz_out": int(st.nnz_out), "bins": int(st.bins),
           "ms_build_device": round(st.ms_build, 3), "ms_symbolic_device": round(st.ms_symbolic, 3),
           "ms_symbolic_pass": round(st.ms_symbolic_pass, 3), "ms_plan_wall": round(st.ms_plan, 3),
           "ms_numeric_device": [round(x, 3) for x in runs]}
    print(json.dumps(out), flush=True)
    plan.free()
    ctx.close()


README_WORKLOAD = ("nips-shape tensor (dims 2482x2862x14036x17, 3101609 unique nnz, Zipf(1.1), seed 2) contracted with "
                   "itself over modes {1,2} (the reference README's 4d.meta): 4-mode output, rows and columns are "
                   "(mode 0, mode 3) pairs, 42194 each")


def readme_bench(P, args):
    """The shape of the reference's one published run (README.md:25-38:
    nips with 4d.meta, c = 2) on the synthetic nips-shape tensor: per-phase
    device times at the reference's boundaries (HT_B build, HT_A build,
    preprocessing, contraction) for each method, beside the oracle's phases
    on the host -- its key tables in full, its contraction on a row sample
    (the synthetic tensor's output is dense-ish, 4.8e8 nonzeros, so the
    full CPU contraction takes minutes)."""
    from oracle import orc

    ctx = P.Context(0)
    h, v = P.gen_coo(SEED, DIMS, NNZ, P.DIST_ZIPF, 1.1)
    A, B = P.self_contraction_inputs(h, v, [1, 2])
    dA, dB = P.DeviceTensor.upload(A, ctx), P.DeviceTensor.upload(B, ctx)
    avail = {"two_pass": P.METHOD_TWO_PASS, "upper_bound": P.METHOD_UPPER_BOUND}
    if P.learned_available():
        avail["learned"] = P.METHOD_LEARNED
    res = {}
    for name, m in avail.items():
        r = run_method(P, ctx, dA, dB, m, max(1, args.steps), args.warmup)
        st = r["stats"]
        res[name] = {"ms_per_step": round(r["ms_per_step"], 3), "ms_build_b": round(st["ms_build_b"], 3),
                     "ms_build_a": round(st["ms_build_a"], 3), "ms_preprocess": round(st["ms_preprocess"], 3),
                     "ms_symbolic": round(st["ms_symbolic"], 3), "ms_numeric": round(st["ms_numeric"], 3),
                     "flops": int(st["flops"]), "nnz_out": int(st["nnz_out"]), "fallback_rows": int(st["fallback_rows"]),
                     "spill_rows": int(st["spill_rows"]), "out_capacity": int(st["out_capacity"])}
        print(f"# readme {name}: {r['ms_per_step']:.2f} ms/step", file=sys.stderr, flush=True)
    F = next(iter(res.values()))["flops"]
    out = {"metric": "useful multiply-adds/s, nips-shape tensor contracted over {1,2} (the reference README's run shape)",
           "config": {"workload": README_WORKLOAD, "flops": F, "nnz_out": next(iter(res.values()))["nnz_out"]},
           "methods": {k: dict(x, MADps=F / (x["ms_per_step"] * 1e-3)) for k, x in res.items()}}
    if not args.no_cpu:
        # the oracle on every 64th row (rows dealt by A's mode-0 and mode-3
        # pair), its contraction time scaled by the flop share
        rows = A.hedges_array[:, 0].astype(np.int64) * 17 + A.hedges_array[:, 3]
        sel = rows % 64 == 0
        As = P.Tensor(A.hedges_array[sel], A.val_array[sel], A.dimensions_array, A.dim_tc, A.num_key_dimensions)
        t0 = time.perf_counter()
        oh, ov, info = orc.tc_fksCuckoo(As, B)
        wall = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "kind": "port", "cores": 1, "unit": "MAD/s", "value": info["flops"] / info["t_contract"],
            "sample": f"every 64th output row (A's mode-0/mode-3 pair index mod 64 == 0): {info['flops']} of {F} "
                      f"multiply-adds; HT_B and HT_A built in full",
            "phases_s": {"ht_b": round(info["t_build_b"], 3), "ht_a_sample": round(info["t_build_a"], 3),
                         "preprocess_sample": round(info["t_preprocess"], 3),
                         "contraction_sample": round(info["t_contract"], 3),
                         "contraction_whole_estimate": round(info["t_contract"] * F / max(info["flops"], 1), 1)},
            "seconds": wall, "nnz_out_sample": int(oh.shape[0])}
    print(json.dumps(out), flush=True)
    ctx.close()


TWO_WORKLOAD = ("two-tensor contraction: A 10000x5000x200 (seed 3), B 5000x200x8000 (seed 4), 5M unique nnz "
                "each, Zipf(1.1) per mode, FP64 U(0,1]; A modes {1,2} with B modes {0,1} -> 10000x8000")


def two_tensor_bench(P, args):
    """Two-tensor configuration: exact two-pass, upper-bound and learned
    single pass (fallback rate from the exact counts), and the FP32 value
    modes with their error against the oracle on 40 sampled output rows."""
    ctx = P.Context(0)
    ha, va = P.gen_coo(3, [10000, 5000, 200], 5_000_000, P.DIST_ZIPF, 1.1)
    hb, vb = P.gen_coo(4, [5000, 200, 8000], 5_000_000, P.DIST_ZIPF, 1.1)
    A = P.make_free_first(P.Tensor(ha, va, None, np.array([1, 2, 0], np.uint32), 2))
    B = P.Tensor(hb, vb, None, np.array([0, 1, 2], np.uint32), 2)
    dA, dB = P.DeviceTensor.upload(A, ctx), P.DeviceTensor.upload(B, ctx)
    avail = {"two_pass": P.METHOD_TWO_PASS, "upper_bound": P.METHOD_UPPER_BOUND}
    if P.learned_available():
        avail["learned"] = P.METHOD_LEARNED
    res = {name: run_method(P, ctx, dA, dB, m, args.steps, args.warmup) for name, m in avail.items()}
    for name, r in res.items():
        print(f"# {name}: {r['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    head = "learned" if "learned" in res else "two_pass"
    st = res[head]["stats"]
    F = int(st["flops"])
    # oracle on 40 sampled rows (A restricted to them, all of B)
    from oracle import orc

    rng = np.random.default_rng(11)
    present = np.unique(ha[:, 0])
    rows = np.sort(rng.choice(present[present >= 100], 40, replace=False))
    sel = np.isin(ha[:, 0], rows)
    Asub = P.Tensor(ha[sel], va[sel], A.dimensions_array, A.dim_tc, A.num_key_dimensions)
    oh, ov, _ = orc.tc_fksCuckoo(Asub, B)
    ref = _sorted(oh, ov)
    fA, fB = P.DeviceTensor.upload_f32(A, ctx), P.DeviceTensor.upload_f32(B, ctx)
    f32 = {}
    for vt_name, vt in (("twosum", P.VAL_F32_TWOSUM), ("f64acc", P.VAL_F32_F64ACC), ("plain", P.VAL_F32_PLAIN)):
        x = run_method(P, ctx, fA, fB, avail[head], args.steps, args.warmup, value_type=vt)
        HB, HA = P.buildHashTable(fB, ctx), P.buildHashTable(fA, ctx)
        O = P.tc_fksCuckoo(HA, HB, download=False, method=avail[head], value_type=vt)
        h = O.download().hedges_array
        hi = O.download_f32().astype(np.float64)
        keep = np.isin(h[:, 0], rows)
        hh, vv = _sorted(h[keep], hi[keep])
        rec = {"ms_per_step": round(x["ms_per_step"], 4), "MADps": F / (x["ms_per_step"] * 1e-3),
               "sample_coords_equal": bool(np.array_equal(hh, ref[0]))}
        if rec["sample_coords_equal"]:
            rel = np.abs(vv - ref[1]) / np.abs(ref[1])
            rec.update(max_rel_err=float(rel.max()), mean_rel_err=float(rel.mean()))
        f32[vt_name] = rec
        O.free()
        print(f"# f32 {vt_name}: {rec['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    r = res[head]
    out = {
        "metric": "useful multiply-adds/s, two-tensor contraction (learned 1-pass vs exact 2-pass)",
        "value": F / (r["ms_per_step"] * 1e-3),
        "unit": "MAD/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator; no dataset)",
        "config": {"workload": TWO_WORKLOAD, "method": head, "flops": F, "nnz_out": int(st["nnz_out"]),
                   "rows": int(st["rows"]),
                   "step": "build HT_B + build HT_A + tc_fksCuckoo, inputs resident in HBM"},
        "methods": {k: {"ms_per_step": round(x["ms_per_step"], 4), "MADps": F / (x["ms_per_step"] * 1e-3),
                        "ms_build_b": round(x["stats"]["ms_build_b"], 4), "ms_build_a": round(x["stats"]["ms_build_a"], 4),
                        "ms_symbolic": round(x["stats"]["ms_symbolic"], 4), "ms_numeric": round(x["stats"]["ms_numeric"], 4),
                        "kernel_numeric_ms": round(x["kernel_ms_per_launch"], 4),
                        "fallback_rows": int(x["stats"]["fallback_rows"]),
                        "fallback_rate": x["stats"]["fallback_rows"] / max(x["stats"]["rows"], 1)}
                    for k, x in res.items()},
        "f32": f32,
        "f32_error_reference": "oracle (FP64, original values) on 40 sampled output rows",
    }
    if "two_pass" in res and head != "two_pass":
        out["speedup_vs_two_pass"] = res["two_pass"]["ms_per_step"] / r["ms_per_step"]
    print(json.dumps(out), flush=True)
    ctx.close()


def chain_bench(P, args):
    """Two-step contraction chain: the nips-shape self-contraction
    O(i, l) = sum A(i, j, k, m) A(l, j, k, m), then O2(i, n) = sum_l O(i, l)
    C(l, n) with C a 2482 x 2000 sparse matrix (seed 5, 500000 nnz, Zipf(1.1)).
    The second step needs O grouped by its row mode i.  "keyed" hands O over
    with that table formed from the output's row grouping
    (ptc_tc_fksCuckoo_keyed); "rebuild" builds it from O's COO, as a chain of
    reference calls would (buildHashTable per contraction).  A step is the
    whole chain from device-resident inputs: both steps' tables and
    contractions."""
    import ctypes as C

    L = P.lib()
    ctx = P.Context(0)
    h, v = P.gen_coo(SEED, DIMS, NNZ, P.DIST_ZIPF, 1.1)
    A, B = P.self_contraction_inputs(h, v, CONTRACTED)
    hc, vc = P.gen_coo(5, [2482, 2000], 500000, P.DIST_ZIPF, 1.1)
    Cm = P.Tensor(hc, vc, None, np.array([0, 1], np.uint32), 1)
    dA, dB, dC = (P.DeviceTensor.upload(T, ctx) for T in (A, B, Cm))
    method = P.METHOD_LEARNED if P.learned_available() else P.METHOD_TWO_PASS
    opts = P.options(method=method)
    st1, st2 = P.Stats(), P.Stats()
    perm = np.array([0, 1], np.uint32)
    pp = P._p32(perm)

    def step(keyed):
        hb, ha, hcm, o, ho, o2 = (C.c_void_p() for _ in range(6))
        fr = C.c_int(0)
        rc = L.ptc_hashtable_build(ctx.handle, dB._h, C.byref(hb))
        rc = rc or L.ptc_hashtable_build(ctx.handle, dA._h, C.byref(ha))
        if keyed:
            rc = rc or L.ptc_tc_fksCuckoo_keyed(ctx.handle, ha, hb, dA._h, dB._h, C.byref(opts), 1, pp, C.byref(o),
                                                C.byref(ho), C.byref(fr), C.byref(st1))
            rc = rc or (0 if fr.value == 1 else -7)
        else:
            rc = rc or L.ptc_tc_fksCuckoo(ctx.handle, ha, hb, dA._h, dB._h, C.byref(opts), C.byref(o), C.byref(st1))
            rc = rc or L.ptc_dtensor_set_keys(o, 1, pp)
            rc = rc or L.ptc_hashtable_build(ctx.handle, o, C.byref(ho))
        rc = rc or L.ptc_hashtable_build(ctx.handle, dC._h, C.byref(hcm))
        rc = rc or L.ptc_tc_fksCuckoo(ctx.handle, ho, hcm, o, dC._h, C.byref(opts), C.byref(o2), C.byref(st2))
        for t in (hb, ha, ho, hcm):
            L.ptc_hashtable_free(t)
        L.ptc_dtensor_free(o2)
        L.ptc_dtensor_free(o)
        if rc != 0:
            raise P.PtcError(rc, ctx.last_error())
        return st2.ms_build_a

    res = {}
    for name, keyed in (("rebuild", False), ("keyed", True)):
        for _ in range(args.warmup):
            step(keyed)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(keyed)
        ctx.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / args.steps
        ctx.set_timing(True)  # the O table's device time, outside the timed steps
        tab = [step(keyed) for _ in range(2)]
        ctx.set_timing(False)
        res[name] = {"ms_per_step": round(ms, 4), "ms_table_O": round(sum(tab) / len(tab), 4)}
        print(f"# {name}: {ms:.3f} ms/step, O table {res[name]['ms_table_O']:.3f} ms", file=sys.stderr, flush=True)
    F = int(st1.flops) + int(st2.flops)
    out = {
        "metric": "useful multiply-adds/s, two-step contraction chain (nips-shape self-contraction, then x matrix)",
        "value": F / (res["keyed"]["ms_per_step"] * 1e-3),
        "unit": "MAD/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["keyed"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator; no dataset)",
        "config": {"workload": WORKLOAD + "; then O x C, C 2482x2000, 500000 nnz, Zipf(1.1), seed 5",
                   "method": "learned" if method == P.METHOD_LEARNED else "two_pass",
                   "flops": F, "nnz_O": int(st1.nnz_out), "nnz_out": int(st2.nnz_out),
                   "step": "tables + contraction of both steps, inputs resident in HBM"},
        "chain": res,
        "speedup_keyed_vs_rebuild": res["rebuild"]["ms_per_step"] / res["keyed"]["ms_per_step"],
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def stream_copy_gbps(mb: int = 1024, reps: int = 10) -> float:
    """Achievable HBM bandwidth on this GPU: a device-to-device copy of `mb`
    MiB (read + write), timed with HIP events; the best of `reps`."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mb << 20
    a, b = C.c_void_p(), C.c_void_p()
    e0, e1 = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(n)) or hip.hipMalloc(C.byref(b), C.c_size_t(n)):
        return 0.0
    hip.hipEventCreate(C.byref(e0))
    hip.hipEventCreate(C.byref(e1))
    hip.hipMemset(a, 1, C.c_size_t(n))
    best = 0.0
    for _ in range(reps):
        hip.hipEventRecord(e0, None)
        hip.hipMemcpyAsync(b, a, C.c_size_t(n), 3, None)
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = C.c_float(0)
        hip.hipEventElapsedTime(C.byref(ms), e0, e1)
        if ms.value > 0:
            best = max(best, 2 * n / (ms.value * 1e-3) / 1e9)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    return best


def load_traffic():
    f = ROOT / "profiles" / "traffic_numeric_pass_nips.json"
    if f.exists():
        try:
            return json.loads(f.read_text())
        except Exception:
            return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--methods", default="two_pass,upper_bound,learned")
    ap.add_argument("--no-f32", action="store_true", help="skip the FP32 compensated leg")
    ap.add_argument("--dense-max", type=int, default=0, help="batchplan: dense LDS column limit (0 = default)")
    ap.add_argument("--lds-hash", type=int, default=0, help="batchplan: LDS hash rows limit (0 = default)")
    ap.add_argument("--chunk", type=int, default=0, help="batchplan: multiply-adds per work item (0 = default)")
    ap.add_argument("--config", default="nips", choices=["nips", "batch", "batchplan", "twotensor", "chain", "readme"],
                    help="nips: the headline self-contraction (default); batch: 256 uneven contractions; "
                         "twotensor: the 5M-nnz two-tensor contraction; chain: two contractions in a row")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        copy_calibration(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the contraction shards by output "
              "row; multi-GPU runs are not part of this build)", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    import partc_amd as P

    if args.config == "batch":
        batch_bench(P, args)
        return
    if args.config == "batchplan":
        batch_plan_bench(P, args)
        return
    if args.config == "readme":
        readme_bench(P, args)
        return
    if args.config == "twotensor":
        two_tensor_bench(P, args)
        return
    if args.config == "chain":
        chain_bench(P, args)
        return

    h, v = P.gen_coo(SEED, DIMS, NNZ, P.DIST_ZIPF, 1.1)
    A, B = P.self_contraction_inputs(h, v, CONTRACTED)
    ctx = P.Context(0)
    dA = P.DeviceTensor.upload(A, ctx)
    dB = P.DeviceTensor.upload(B, ctx)

    avail = {"two_pass": P.METHOD_TWO_PASS, "upper_bound": P.METHOD_UPPER_BOUND}
    try:
        if P.learned_available():
            avail["learned"] = P.METHOD_LEARNED
    except AttributeError:
        pass
    results = {}
    for name in args.methods.split(","):
        if name in avail:
            results[name] = run_method(P, ctx, dA, dB, avail[name], args.steps, args.warmup)
            print(f"# {name}: {results[name]['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)

    head = "learned" if "learned" in results else ("two_pass" if "two_pass" in results else next(iter(results)))
    r = results[head]
    st = r["stats"]
    F = int(st["flops"])
    nnz_c = int(st["nnz_out"])
    d, c = 4, 3
    d_o = (d - c) * 2
    value = F / (r["ms_per_step"] * 1e-3)
    alg = algorithmic_bytes(NNZ, d, c, d, F, nnz_c, d_o)
    kb = kernel_bytes(NNZ, F, d, c, nnz_c, d_o)
    kms = r["kernel_ms_per_launch"]
    achieved = kb / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    traffic = load_traffic()
    roof = {
        "kernel": "numeric row pass: k_pass<NUMERIC,dense> row launch + its merge launch",
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic.get("bytes_per_launch") if traffic else None,
        "algorithmic_bytes_per_launch": int(kb),
        "kernel_ms": round(kms, 4),
    }
    copy_gbps = stream_copy_gbps()
    sb = streamed_bytes(NNZ, F, nnz_c, d_o)
    streamed = sb / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    roof["streamed_bytes_per_launch"] = int(sb)
    roof["streamed_GBps"] = round(streamed, 1)
    if traffic and kms > 0:  # the HBM bytes the counters measured, over the same live kernel time
        tgbps = traffic["bytes_per_launch"] / (kms * 1e-3) / 1e9
        roof["traffic_GBps"] = round(tgbps, 1)
        roof["traffic_frac"] = round(tgbps / HBM_PEAK_GBS, 4)
    if copy_gbps > 0:  # what a plain device copy reaches on this GPU
        roof["achievable_copy"] = round(copy_gbps, 1)
        roof["frac_of_achievable"] = round(achieved / copy_gbps, 4)
        roof["streamed_frac_of_achievable"] = round(streamed / copy_gbps, 4)
    out = {
        "metric": "useful multiply-adds/s, nips-shape 4-mode self-contraction (learned 1-pass vs exact 2-pass)",
        "value": value,
        "unit": "MAD/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator; no dataset)",
        "config": {
            "workload": WORKLOAD,
            "method": head,
            "flops": F,
            "nnz_out": nnz_c,
            "step": "build HT_B + build HT_A + tc_fksCuckoo, inputs resident in HBM",
        },
        "step_hbm": {
            "algorithmic_bytes": int(alg),
            "GBps": round(alg / (r["ms_per_step"] * 1e-3) / 1e9, 1),
            "frac": round(alg / (r["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        },
        "methods": {
            k: {
                "ms_per_step": round(x["ms_per_step"], 4),
                "MADps": F / (x["ms_per_step"] * 1e-3),
                "ms_build_b": round(x["stats"]["ms_build_b"], 4),
                "ms_build_a": round(x["stats"]["ms_build_a"], 4),
                "ms_preprocess": round(x["stats"]["ms_preprocess"], 4),
                "ms_symbolic": round(x["stats"]["ms_symbolic"], 4),
                "ms_numeric": round(x["stats"]["ms_numeric"], 4),
                "ms_contract": round(x["stats"]["ms_total"], 4),
                "kernel_numeric_ms": round(x["kernel_ms_per_launch"], 4),
                "kernel_symbolic_ms": round(x["stats"]["ms_kernel_symbolic"], 4),
                "fallback_rows": int(x["stats"]["fallback_rows"]),
                "heavy_rows": int(x["stats"]["heavy_rows"]),
                "out_capacity": int(x["stats"]["out_capacity"]),
                "device_peak_bytes": int(x["stats"]["device_peak_bytes"]),
            }
            for k, x in results.items()
        },
        "roofline": roof,
    }
    if "two_pass" in results and head != "two_pass":
        out["speedup_vs_two_pass"] = results["two_pass"]["ms_per_step"] / r["ms_per_step"]
    ref = None
    if not args.no_cpu and rank == 0:
        out["cpu_baseline"], ref = cpu_baseline(P, h, v)
        cb = out["cpu_baseline"]
        if cb["nnz_out"] != nnz_c:
            out["cpu_baseline"]["note"] = f"nnz mismatch: cpu {cb['nnz_out']} gpu {nnz_c}"
        cb["c1"]["gpu_two_pass"] = gpu_c1(P, ctx)
    if not args.no_f32:
        # the same tensors with values rounded to FP32: two-sum compensated
        # FP32 accumulation, exact products accumulated in FP64, and the plain
        # FP32 path for comparison
        fA = P.DeviceTensor.upload_f32(A, ctx)
        fB = P.DeviceTensor.upload_f32(B, ctx)
        f32 = {}
        for name in args.methods.split(","):
            if name not in avail:
                continue
            for vt_name, vt in (("twosum", P.VAL_F32_TWOSUM), ("f64acc", P.VAL_F32_F64ACC),
                                ("plain", P.VAL_F32_PLAIN)):
                x = run_method(P, ctx, fA, fB, avail[name], args.steps, args.warmup, value_type=vt)
                rec = {"ms_per_step": round(x["ms_per_step"], 4), "MADps": F / (x["ms_per_step"] * 1e-3),
                       "kernel_numeric_ms": round(x["kernel_ms_per_launch"], 4),
                       "fallback_rows": int(x["stats"]["fallback_rows"])}
                if ref is not None:
                    rec.update(f32_errors(P, ctx, fA, fB, avail[name], vt, ref))
                f32[f"{name}_{vt_name}"] = rec
                print(f"# f32 {name} {vt_name}: {rec['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
        out["f32"] = f32
    if rank == 0:
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
