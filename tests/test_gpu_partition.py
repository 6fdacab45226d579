"""Partitioned grouping of key tables (PTC_GROUP_PARTITION: nonzeros
partitioned into buckets by key hash, each bucket grouped in an LDS table,
buckets larger than the LDS table grouped in their own region of a global
table) against the definition of a key table and the oracle: every key's
slice holds exactly the ids with that key, absent keys miss, and
contractions through partitioned tables equal the oracle's (coordinates
bit-exact after sorting, values within 1e-12 relative)."""
import numpy as np
import pytest

import partc_amd as P
from oracle import orc
from tests.helpers import golden_cases, max_rel, random_case, sort_coo, tc_inputs

pytestmark = pytest.mark.gpu
TOL = 1e-12


def key_groups(T):
    """{key tuple: ascending ids} of a host tensor's key modes."""
    modes = T.dim_tc[: T.num_key_dimensions]
    keys = T.hedges_array[:, modes]
    order = np.lexsort(keys.T[::-1])
    groups = {}
    for i in order:
        groups.setdefault(tuple(keys[i].tolist()), []).append(int(i))
    return groups


def meta(h, v, key_modes):
    d = h.shape[1]
    perm = list(key_modes) + [m for m in range(d) if m not in key_modes]
    return P.Tensor(h, v, None, np.array(perm, np.uint32), len(key_modes))


SHAPES = {
    # wide uniform keys: every bucket in LDS
    "uniform": lambda: meta(*P.gen_coo(91, [50, 400, 300, 30], 30000, P.DIST_UNIFORM), [1, 2, 3]),
    # one key holding about a quarter of 60000 nonzeros: a bucket far larger
    # than the LDS table with few distinct keys
    "hot-key": lambda: meta(*P.gen_coo(94, [30, 4000, 5000, 60], 60000, P.DIST_ZIPF, 2.0), [1, 2, 3]),
    # 12288 nonzeros with distinct keys in 4 buckets of about 3072: buckets
    # whose keys outgrow the LDS table are grouped in their global region
    "global-region": lambda: meta(*P.gen_coo(98, [50, 4000, 5000, 60], 12288, P.DIST_UNIFORM), [1, 2, 3]),
    # fewer than 2048 nonzeros: one bucket
    "one-bucket": lambda: meta(*P.gen_coo(95, [40, 300, 200], 1500, P.DIST_UNIFORM), [1, 2]),
    # a non-key tuple wider than 32 bits (64-bit column keys)
    "rest64": lambda: meta(*P.gen_coo(96, [70000, 70000, 70000, 70000], 20000, P.DIST_UNIFORM), [0, 1]),
    # a direct (small) key space: one counting pass, every key its own bucket
    "direct": lambda: meta(*P.gen_coo(92, [700, 60, 50], 20000, P.DIST_ZIPF), [0]),
    # the largest such key space (4096 keys), hot rows, several tiles
    "direct-4096": lambda: meta(*P.gen_coo(99, [4096, 300, 200], 200000, P.DIST_ZIPF, 1.3), [0]),
    # a direct key space over two modes with a 64-bit non-key tuple
    "direct-rest64": lambda: meta(*P.gen_coo(100, [40, 50, 90000, 90000], 30000, P.DIST_UNIFORM), [0, 1]),
}


@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_partitioned_table_slices(ctx, shape):
    print(f"[case] partitioned slices {shape}", flush=True)
    T = SHAPES[shape]()
    groups = key_groups(T)
    HP = P.buildHashTable(T, ctx, grouping=P.GROUP_PARTITION)
    assert HP.info() == (len(groups), T.N)
    keys = np.array(list(groups.keys()), dtype=np.uint32)
    for want, have in zip(groups.values(), HP.find(keys)):
        assert sorted(have.tolist()) == want
    miss = np.full((3, keys.shape[1]), 7, np.uint32)
    miss[:, 0] = np.iinfo(np.uint32).max - 1
    assert not HP.find(miss, with_ids=False).any()


@pytest.mark.parametrize("name", sorted(golden_cases()))
def test_partitioned_grouping_golden(ctx, name):
    case = golden_cases()[name]
    A, B = tc_inputs(case)
    O = P.tc_fksCuckoo(P.buildHashTable(A, ctx, grouping=P.GROUP_PARTITION),
                       P.buildHashTable(B, ctx, grouping=P.GROUP_PARTITION))
    h, v = sort_coo(O.hedges_array, O.val_array)
    assert np.array_equal(h, case["coords"])
    assert max_rel(v, case["vals_ordered"]) <= TOL


@pytest.mark.parametrize("i", range(4))
def test_partitioned_grouping_random(ctx, i):
    shapes = [(3, [200, 300, 400], 50000, [1, 2], False), (4, [60, 70, 80, 9], 40000, [1, 2, 3], True),
              (3, [900, 40, 30], 30000, [1, 2], True), (4, [40, 3000, 4000, 50], 60000, [1, 2, 3], True)]
    d, dims, nnz, contracted, zipf = shapes[i]
    A, B = random_case(np.random.default_rng(i), d, dims, nnz, contracted, zipf, seed=800 + i)
    print(f"[case] partitioned grouping random {i}", flush=True)
    oh, ov, _ = orc.tc_fksCuckoo(A, B)
    oh, ov = sort_coo(oh, ov)
    for method in (P.METHOD_TWO_PASS, P.METHOD_UPPER_BOUND):
        O = P.tc_fksCuckoo(P.buildHashTable(A, ctx, grouping=P.GROUP_PARTITION),
                           P.buildHashTable(B, ctx, grouping=P.GROUP_PARTITION), method=method)
        h, v = sort_coo(O.hedges_array, O.val_array)
        assert np.array_equal(h, oh)
        assert max_rel(v, ov) <= TOL


def test_partitioned_fp32_matches_hashed(ctx):
    """FP32 tables (8-byte payloads) group the same way."""
    print("[case] partitioned fp32", flush=True)
    h, v = P.gen_coo(97, [80, 500, 600, 20], 40000, P.DIST_ZIPF)
    A, B = P.self_contraction_inputs(h, v, [1, 2, 3])
    dA, dB = P.DeviceTensor.upload_f32(A, ctx), P.DeviceTensor.upload_f32(B, ctx)
    outs = []
    for g in (P.GROUP_PARTITION, P.GROUP_HASH):
        O = P.tc_fksCuckoo(P.buildHashTable(dA, ctx, grouping=g), P.buildHashTable(dB, ctx, grouping=g),
                           download=False, value_type=P.VAL_F32_F64ACC)
        t = O.download()
        outs.append(sort_coo(t.hedges_array, t.val_array))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert max_rel(outs[0][1], outs[1][1]) <= 1e-6


def test_auto_grouping_counts_small_key_spaces(ctx):
    """The nips-shape A table (2482 rows) is grouped by one counting pass
    by default: the same slices and ids as the stable sort grouping, and the
    same row of every position (slice indices)."""
    print("[case] nips-shape A counted", flush=True)
    h, v = P.gen_coo(2, [2482, 2862, 14036, 17], 3101609, P.DIST_ZIPF, 1.1)
    A, B = P.self_contraction_inputs(h, v, [1, 2, 3])
    dA = P.DeviceTensor.upload(A, ctx)
    HC = P.buildHashTable(dA, ctx)
    HS = P.buildHashTable(dA, ctx, grouping=P.GROUP_SORT)
    assert HC.info() == HS.info()
    keys = np.unique(A.hedges_array[:, [0]], axis=0)
    for x, y in zip(HC.find(keys), HS.find(keys)):
        assert np.array_equal(np.sort(x), y)


def test_auto_grouping_partitions_wide_keys(ctx):
    """The nips-shape B table (30-bit keys) is grouped by partitioning by
    default, and its slices equal the single-table hash grouping's."""
    print("[case] nips-shape B partitioned", flush=True)
    h, v = P.gen_coo(2, [2482, 2862, 14036, 17], 3101609, P.DIST_ZIPF, 1.1)
    A, B = P.self_contraction_inputs(h, v, [1, 2, 3])
    dB = P.DeviceTensor.upload(B, ctx)
    HA = P.buildHashTable(dB, ctx)
    HH = P.buildHashTable(dB, ctx, grouping=P.GROUP_HASH)
    assert HA.info() == HH.info()
    rng = np.random.default_rng(5)
    keys = B.hedges_array[rng.choice(B.N, 2000, replace=False)][:, [1, 2, 3]]
    a, b = HA.find(keys), HH.find(keys)
    for x, y in zip(a, b):
        assert np.array_equal(np.sort(x), np.sort(y))
