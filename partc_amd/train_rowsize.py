"""Train the row-size network on exact symbolic-pass counts.

Training data (all seeded, generated by the engine's generator):
  * 64 two-tensor contractions of one family: A 3-mode 10000x5000x200,
    B 3-mode 5000x200x8000, Zipf(1.1) per mode, contract A{1,2} with B{0,1};
    nnz per tensor log-uniform in [2e5, 5e6] (the family's full size is 5e6);
  * 16 nips-shape 4-mode self-contractions over {1,2,3} with nnz log-uniform
    in [3e5, 3.1e6];
  * 32 jobs of the uneven batch family (3/4 modes, dims log-uniform in
    [64, 20000], nnz log-uniform in [1e3, 1e6], all modes but 0 contracted).
Each contraction runs the engine's preprocessing and exact symbolic pass on
the GPU (ptc_row_profile) and contributes one (features, exact nnz) sample per
output row.  The network is trained on the GPU by the library's own forward /
backward / Adam kernels; held-out contractions (other seeds) measure the
fallback rate at the capacity margin the LEARNED method uses.

Run:  python -m partc_amd.train_rowsize [--out partc_amd/models/rowsize_mlp.bin]
"""
from __future__ import annotations

import argparse
import json
import math
import multiprocessing as mp
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def family_specs(split: str):
    """(kind, params) of every training / held-out contraction."""
    base = {"train": 0, "heldout": 10_000}[split]
    n3 = 64 if split == "train" else 8
    n2 = 16 if split == "train" else 4
    n4 = 32 if split == "train" else 8
    specs = []
    rng = np.random.default_rng(123 + base)
    for i in range(n3):
        nnz = int(math.exp(rng.uniform(math.log(2e5), math.log(5e6))))
        specs.append(("two", dict(seed_a=30_000 + base + 2 * i, seed_b=30_001 + base + 2 * i, nnz=nnz)))
    for i in range(n2):
        nnz = int(math.exp(rng.uniform(math.log(3e5), math.log(3.1e6))))
        specs.append(("nips", dict(seed=50_000 + base + i, nnz=nnz)))
    for i in range(n4):
        specs.append(("job", dict(seed=70_000 + base + i)))
    return specs


def make_inputs(spec):
    """Generates one contraction's inputs (host arrays) in a worker."""
    import partc_amd as P

    kind, p = spec
    if kind == "two":
        ha, va = P.gen_coo(p["seed_a"], [10000, 5000, 200], p["nnz"], P.DIST_ZIPF)
        hb, vb = P.gen_coo(p["seed_b"], [5000, 200, 8000], p["nnz"], P.DIST_ZIPF)
        return kind, (ha, va, [1, 2, 0], 2), (hb, vb, [0, 1, 2], 2)
    if kind == "nips":
        h, v = P.gen_coo(p["seed"], [2482, 2862, 14036, 17], p["nnz"], P.DIST_ZIPF)
        return kind, (h, v, [1, 2, 3, 0], 3), None
    d, dims, nnz = P.gen_job_shape(p["seed"])
    h, v = P.gen_coo(p["seed"], dims.tolist(), nnz, P.DIST_ZIPF)
    perm = list(range(1, d)) + [0]
    return kind, (h, v, perm, d - 1), None


def profile(P, ctx, item):
    kind, a, b = item
    ha, va, pa, ca = a
    A = P.Tensor(ha, va, None, np.array(pa, np.uint32), ca)
    B = P.get_sparse_tensor(A) if b is None else P.Tensor(b[0], b[1], None, np.array(b[2], np.uint32), b[3])
    HB = P.buildHashTable(B, ctx)
    HA = P.buildHashTable(P.make_free_first(A), ctx)
    X, ub, nnz = P.row_profile(HA, HB)
    HA.free()
    HB.free()
    keep = ub >= 1
    return X[keep], ub[keep], nnz[keep].astype(np.float32)


def collect(P, ctx, split, pool):
    specs = family_specs(split)
    Xs, us, ys = [], [], []
    t0 = time.time()
    for i, item in enumerate(pool.imap(make_inputs, specs)):
        X, ub, y = profile(P, ctx, item)
        Xs.append(X)
        us.append(ub)
        ys.append(y)
        if (i + 1) % 8 == 0:
            print(f"[{split}] {i + 1}/{len(specs)} contractions profiled ({time.time() - t0:.1f}s)", flush=True)
    return np.concatenate(Xs), np.concatenate(us), np.concatenate(ys)


def evaluate(model, X, ub, y, scale):
    pred = model.predict(X, ub)
    cap = np.minimum(np.ceil(pred * scale), ub)
    return {
        "rows": int(y.shape[0]),
        "fallback_rate": float(np.mean(cap < y)),
        "capacity_over_exact": float(cap.sum() / y.sum()),
        "upper_bound_over_exact": float(ub.sum() / y.sum()),
        "median_rel_err": float(np.median(np.abs(pred - y) / y)),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(ROOT / "partc_amd" / "models" / "rowsize_mlp.bin"))
    ap.add_argument("--report", default="")
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--scale", type=float, default=1.15)
    ap.add_argument("--workers", type=int, default=12)
    args = ap.parse_args()

    import partc_amd as P

    # The generator workers are forked before this process touches the GPU.
    pool = mp.get_context("fork").Pool(args.workers)
    ctx = P.Context(0)
    Xtr, utr, ytr = collect(P, ctx, "train", pool)
    Xte, ute, yte = collect(P, ctx, "heldout", pool)
    pool.close()
    pool.join()
    print(f"train rows {ytr.shape[0]}, held-out rows {yte.shape[0]}", flush=True)
    model = P.RowSizeModel(ctx, seed=1)
    before = evaluate(model, Xte, ute, yte, args.scale)
    t0 = time.time()
    losses = model.train(Xtr, utr, ytr, epochs=args.epochs, batch=args.batch, lr=args.lr)
    t_train = time.time() - t0
    after = evaluate(model, Xte, ute, yte, args.scale)
    model.save(args.out)
    rep = {
        "train_rows": int(ytr.shape[0]),
        "heldout_rows": int(yte.shape[0]),
        "epochs": args.epochs,
        "batch": args.batch,
        "lr": args.lr,
        "train_seconds": t_train,
        "loss_first_epoch": float(losses[0]),
        "loss_last_epoch": float(losses[-1]),
        "heldout_untrained": before,
        "heldout_trained": after,
        "capacity_scale": args.scale,
        "model": str(Path(args.out).relative_to(ROOT)) if Path(args.out).is_relative_to(ROOT) else args.out,
    }
    print(json.dumps(rep, indent=1), flush=True)
    if args.report:
        Path(args.report).write_text(json.dumps(rep, indent=1))


if __name__ == "__main__":
    main()
