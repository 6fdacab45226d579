"""Writes tests/golden/contractions.npz: small contractions with expected
outputs computed here by two formulations that share nothing with the oracle
or the engine:

  * dense: the tensors are densified, permuted so A's contracted modes are
    last and B's first, reshaped to matrices and multiplied with numpy;
    the output coordinate set is the structural (boolean) product.
  * ordered: a plain loop over A's nonzeros in input order that adds
    A.val[a] * B.val[b] into a dict per output coordinate -- the
    summation order the reference's row loop produces (the SPA is seeded by
    the row's first nonzero, fksCuckooTC.hpp:262-290, and later nonzeros are
    added in input order, :294-337, with the product rounded before the add,
    :313,318).  Python floats are IEEE doubles, so these values are exact
    reference values.

Inputs come from the engine's seeded generator (partc_amd.gen_coo) when the
library is built, else from numpy (the fixture stores the inputs, so either
is fine).  Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def unique_coo(rng, dims, nnz, zipf=False):
    d = len(dims)
    seen = set()
    out = []
    while len(out) < nnz:
        if zipf:
            t = tuple(int(min(rng.zipf(1.1) - 1, dims[j] - 1)) for j in range(d))
        else:
            t = tuple(int(rng.integers(0, dims[j])) for j in range(d))
        if t not in seen:
            seen.add(t)
            out.append(t)
    h = np.array(out, dtype=np.uint32).reshape(-1, d)
    v = rng.random(nnz) * (1 - 2 ** -53) + 2 ** -53
    return h, v.astype(np.float64)


def expected(Ah, Av, Aperm, c, Bh, Bv, Bperm, Adims, Bdims):
    """Both formulations. Aperm/Bperm: .meta permutations (contracted first)."""
    dA, dB = Ah.shape[1], Bh.shape[1]
    ca, fa = list(Aperm[:c]), list(Aperm[c:])
    cb, fb = list(Bperm[:c]), list(Bperm[c:])
    # dense
    DA = np.zeros(Adims)
    DA[tuple(Ah[:, j] for j in range(dA))] = Av
    DB = np.zeros(Bdims)
    DB[tuple(Bh[:, j] for j in range(dB))] = Bv
    SA = np.zeros(Adims, dtype=np.int64)
    SA[tuple(Ah[:, j] for j in range(dA))] = 1
    SB = np.zeros(Bdims, dtype=np.int64)
    SB[tuple(Bh[:, j] for j in range(dB))] = 1
    # paired contracted extents must agree: pad to the max
    kdims = [max(Adims[a], Bdims[b]) for a, b in zip(ca, cb)]

    def pad(D, mode_list, ext):
        pads = [(0, 0)] * D.ndim
        for m, e in zip(mode_list, ext):
            pads[m] = (0, e - D.shape[m])
        return np.pad(D, pads)

    DA, SA = pad(DA, ca, kdims), pad(SA, ca, kdims)
    DB, SB = pad(DB, cb, kdims), pad(SB, cb, kdims)
    MA = np.transpose(DA, fa + ca).reshape(int(np.prod([Adims[m] for m in fa])), -1)
    MB = np.transpose(DB, cb + fb).reshape(-1, int(np.prod([Bdims[m] for m in fb])))
    NA = np.transpose(SA, fa + ca).reshape(MA.shape)
    NB = np.transpose(SB, cb + fb).reshape(MB.shape)
    Cd = MA @ MB
    pat = (NA @ NB) > 0
    rows, cols = np.nonzero(pat)
    rc = np.array(np.unravel_index(rows, [Adims[m] for m in fa])).T.reshape(-1, len(fa))
    cc = np.array(np.unravel_index(cols, [Bdims[m] for m in fb])).T.reshape(-1, len(fb))
    coords = np.concatenate([rc, cc], axis=1).astype(np.uint32)
    vals_dense = Cd[rows, cols]
    # ordered (reference summation order)
    bmap = {}
    for b in range(Bh.shape[0]):
        bmap.setdefault(tuple(int(Bh[b, m]) for m in cb), []).append(b)
    acc = {}
    for a in range(Ah.shape[0]):
        key = tuple(int(Ah[a, m]) for m in ca)
        for b in bmap.get(key, ()):
            out = tuple(int(Ah[a, m]) for m in fa) + tuple(int(Bh[b, m]) for m in fb)
            prod = float(Av[a]) * float(Bv[b])
            acc[out] = prod if out not in acc else acc[out] + prod
    order = np.lexsort(coords.T[::-1])
    coords, vals_dense = coords[order], vals_dense[order]
    vals_ordered = np.array([acc[tuple(int(x) for x in t)] for t in coords], dtype=np.float64)
    flops = sum(len(bmap.get(tuple(int(Ah[a, m]) for m in ca), ())) for a in range(Ah.shape[0]))
    return coords, vals_dense, vals_ordered, flops


CASES = [
    # name, A dims, A nnz, zipf, A perm (meta), c, B spec ("self" or (dims, nnz, zipf, perm))
    ("self3_c2", [6, 7, 8], 80, False, [1, 2, 0], 2, "self"),
    ("self4_c3_zipf", [5, 4, 6, 3], 100, True, [1, 2, 3, 0], 3, "self"),
    ("self4_c1", [6, 5, 4, 3], 90, False, [2, 0, 1, 3], 1, "self"),
    ("matrix", [10, 12], 40, False, [1, 0], 1, "self"),
    ("two3", [8, 6, 5], 70, False, [1, 2, 0], 2, ([6, 5, 9], 90, False, [0, 1, 2])),
    ("two_perm", [7, 4, 5], 60, True, [2, 0, 1], 2, ([6, 5, 7], 80, False, [1, 0, 2])),
    ("two_free2", [4, 5, 3, 6], 70, False, [1, 2, 0, 3], 2, ([5, 3, 4, 2], 60, False, [0, 1, 3, 2])),
    ("sparse_rows", [30, 40, 50], 25, False, [1, 2, 0], 2, "self"),
]


def main():
    rng = np.random.default_rng(20261203)
    try:
        sys.path.insert(0, str(HERE.parents[1]))
        import partc_amd as P

        P.lib()
        gen = P
    except Exception:
        gen = None
    out = {}
    for i, (name, dims, nnz, zipf, perm, c, bspec) in enumerate(CASES):
        if gen is not None:
            Ah, Av = gen.gen_coo(100 + i, dims, nnz, gen.DIST_ZIPF if zipf else gen.DIST_UNIFORM)
        else:
            Ah, Av = unique_coo(rng, dims, nnz, zipf)
        if bspec == "self":
            Bh, Bv, bdims, bperm = Ah.copy(), Av.copy(), list(dims), list(perm)
        else:
            bdims, bnnz, bz, bperm = bspec
            if gen is not None:
                Bh, Bv = gen.gen_coo(200 + i, bdims, bnnz, gen.DIST_ZIPF if bz else gen.DIST_UNIFORM)
            else:
                Bh, Bv = unique_coo(rng, bdims, bnnz, bz)
        coords, vd, vo, flops = expected(Ah, Av, np.array(perm), c, Bh, Bv, np.array(bperm), list(dims), list(bdims))
        for k, val in dict(Ah=Ah, Av=Av, Aperm=np.array(perm, np.uint32), Adims=np.array(dims, np.uint32),
                           Bh=Bh, Bv=Bv, Bperm=np.array(bperm, np.uint32), Bdims=np.array(bdims, np.uint32),
                           c=np.array([c], np.uint32), coords=coords, vals_dense=vd, vals_ordered=vo,
                           flops=np.array([flops], np.uint64)).items():
            out[f"{name}/{k}"] = val
        print(f"{name}: nnzA={Ah.shape[0]} nnzB={Bh.shape[0]} out={coords.shape[0]} flops={flops}")
    np.savez_compressed(HERE / "contractions.npz", **out)


if __name__ == "__main__":
    main()
