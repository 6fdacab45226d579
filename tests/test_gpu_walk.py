"""The row passes' walk from flops to A positions, against the oracle.

A row pass expands a work item's flat flop range in groups of 64 flops per
wave; a lane finds its A position from per-wave boundary flags (where each
staged position starts) instead of a search.  These cases put the
boundaries where that walk can go wrong: B slices of length 1 (every lane of
a group in its own position), lengths one either side of multiples of 64
(boundaries on and next to group edges), one long slice (no boundary),
rows of more A positions than one staging batch holds, chunks that start
in the middle of a position, and every workgroup size.  Coordinates
bit-exact after sorting, values within 1e-12 relative.
"""
import numpy as np
import pytest

import partc_amd as P
from oracle import orc
from tests.helpers import max_rel, sort_coo

pytestmark = pytest.mark.gpu
TOL = 1e-12
METHODS = {"two_pass": P.METHOD_TWO_PASS, "upper_bound": P.METHOD_UPPER_BOUND}


def matrix_pair(slice_lens, a_rows, ncols, seed):
    """A(i, j) free-first and B(j, k) contracted-first, 2-mode each.  B key j
    has slice_lens[j] entries (distinct columns); a_rows[i] lists row i's
    contracted keys j."""
    rng = np.random.default_rng(seed)
    bh, bv = [], []
    for j, n in enumerate(slice_lens):
        cols = rng.choice(ncols, size=n, replace=False)
        bh += [(j, int(k)) for k in cols]
        bv += list(rng.uniform(0.5, 1.0, size=n))
    ah, av = [], []
    for i, js in enumerate(a_rows):
        ah += [(i, int(j)) for j in js]
        av += list(rng.uniform(0.5, 1.0, size=len(js)))
    A = P.Tensor(np.array(ah, np.uint32), np.array(av), None, np.array([0, 1], np.uint32), 1)
    B = P.Tensor(np.array(bh, np.uint32), np.array(bv), None, np.array([0, 1], np.uint32), 1)
    return A, B


def check(ctx, A, B, label, **opts):
    print(f"[case] walk {label} opts={opts}", flush=True)
    HB = P.buildHashTable(B, ctx)
    HA = P.buildHashTable(A, ctx)
    st = P.Stats()
    O = P.tc_fksCuckoo(HA, HB, stats=st, **opts)
    h, v = sort_coo(O.hedges_array, O.val_array)
    oh, ov, info = orc.tc_fksCuckoo(A, B)
    oh, ov = sort_coo(oh, ov)
    assert st.flops == info["flops"]
    assert h.shape == oh.shape, (h.shape, oh.shape)
    assert np.array_equal(h, oh)
    assert max_rel(v, ov) <= TOL
    return st


def edge_lengths():
    """1s, then lengths one either side of 64, 128 and 192, then a long one."""
    lens = [1] * 300
    for m in (64, 128, 192):
        lens += [m - 1, m, m + 1] * 4
    lens += [1, 2, 3] * 50 + [1500]
    return lens


@pytest.fixture(scope="module")
def walk_case():
    lens = edge_lengths()
    nk = len(lens)
    rng = np.random.default_rng(71)
    rows = [
        np.arange(nk),                                     # every key: all boundary kinds in one row
        rng.permutation(nk)[:40],                          # a short row
        np.arange(300),                                    # only length-1 slices
        np.array([nk - 1]),                                # one long slice
        np.array([], dtype=np.int64),                      # an empty row
    ]
    rows += [rng.choice(nk, size=int(rng.integers(1, 200)), replace=False) for _ in range(60)]
    return matrix_pair(lens, rows, 3000, seed=72)


@pytest.mark.parametrize("acc", ["auto", "hash"])
@pytest.mark.parametrize("threads", [64, 512, 1024])
@pytest.mark.parametrize("method", sorted(METHODS))
def test_walk_boundaries(ctx, walk_case, method, threads, acc):
    """Dense (direct-mapped) and LDS-hash accumulators walk the same way."""
    A, B = walk_case
    check(ctx, A, B, f"boundaries t={threads} acc={acc}", method=METHODS[method], block_threads=threads,
          acc_mode=P.ACC_AUTO if acc == "auto" else P.ACC_HASH)


@pytest.mark.parametrize("chunk", [64, 100, 127, 4096])
def test_walk_chunks_start_inside_positions(ctx, walk_case, chunk):
    """Small chunks split the rows so chunks begin inside a position."""
    A, B = walk_case
    st = check(ctx, A, B, f"chunk={chunk}", method=P.METHOD_TWO_PASS, chunk_flops=chunk)
    assert st.heavy_rows > 0


@pytest.mark.parametrize("threads", [64, 1024])
def test_walk_rows_longer_than_a_staging_batch(ctx, threads):
    """Rows of 5000 A positions (a staging batch holds 2 per thread), B slices
    of lengths 1..3: many batches per row, boundaries in nearly every group."""
    rng = np.random.default_rng(73)
    lens = list(rng.integers(1, 4, size=6000))
    rows = [rng.choice(6000, size=5000, replace=False) for _ in range(3)] + [np.arange(6000)]
    A, B = matrix_pair(lens, rows, 500, seed=74)
    check(ctx, A, B, f"long rows t={threads}", method=P.METHOD_UPPER_BOUND, block_threads=threads)
    check(ctx, A, B, f"long rows t={threads} chunk", method=P.METHOD_TWO_PASS, block_threads=threads,
          chunk_flops=1000)
