"""Sort grouping of key tables (PTC_GROUP_SORT: stable LSD radix sort, slices
from runs of equal keys) against hash grouping and the oracle: the same
slices, ids ascending inside every slice (the sort is stable), and
contractions through sorted tables equal to the oracle."""
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


@pytest.mark.parametrize("shape", ["hashed", "direct", "zipf-hot"])
def test_sorted_table_slices(ctx, shape):
    if shape == "hashed":
        h, v = P.gen_coo(91, [50, 400, 300, 30], 30000, P.DIST_UNIFORM)
        A, B = P.self_contraction_inputs(h, v, [1, 2, 3])
        T = B
    elif shape == "direct":
        h, v = P.gen_coo(92, [700, 60, 50], 20000, P.DIST_ZIPF)
        A, B = P.self_contraction_inputs(h, v, [1, 2])
        T = A  # keyed by mode 0: a small key space
    else:
        h, v = P.gen_coo(93, [30, 40, 50], 25000, P.DIST_ZIPF, 1.6)  # a few very long slices
        A, B = P.self_contraction_inputs(h, v, [1, 2])
        T = B
    print(f"[case] sorted slices {shape}", flush=True)
    groups = key_groups(T)
    HS = P.buildHashTable(T, ctx, grouping=P.GROUP_SORT)
    HH = P.buildHashTable(T, ctx, grouping=P.GROUP_HASH)
    assert HS.info()[0] == HH.info()[0] == len(groups)
    keys = np.array(list(groups.keys()), dtype=np.uint32)
    got = HS.find(keys)
    for want, have in zip(groups.values(), got):
        assert have.tolist() == want  # exact ids, ascending: stable
    assert np.array_equal(HS.find(keys, with_ids=False), HH.find(keys, with_ids=False))
    for want, have in zip(groups.values(), HH.find(keys)):
        assert sorted(have.tolist()) == want
    # keys that are absent miss
    miss = np.full((3, keys.shape[1]), 7, np.uint32)
    miss[:, 0] = np.iinfo(np.uint32).max - 1
    assert not HS.find(miss, with_ids=False).any()


@pytest.mark.parametrize("name", sorted(golden_cases()))
def test_sorted_grouping_golden(ctx, name):
    case = golden_cases()[name]
    A, B = tc_inputs(case)
    O = P.tc_fksCuckoo(P.buildHashTable(A, ctx, grouping=P.GROUP_SORT), P.buildHashTable(B, ctx, grouping=P.GROUP_SORT))
    h, v = sort_coo(O.hedges_array, O.val_array)
    assert np.array_equal(h, case["coords"])
    assert max_rel(v, case["vals_ordered"]) <= TOL


@pytest.mark.parametrize("i", range(4))
def test_sorted_grouping_random(ctx, i):
    shapes = [(3, [200, 300, 400], 50000, [1, 2], False), (4, [60, 70, 80, 9], 40000, [1, 2, 3], True),
              (3, [900, 40, 30], 30000, [1, 2], True), (3, [80, 3000, 4000], 30000, [0], True)]
    d, dims, nnz, contracted, zipf = shapes[i]
    A, B = random_case(np.random.default_rng(i), d, dims, nnz, contracted, zipf, seed=700 + i)
    print(f"[case] sorted grouping random {i}", flush=True)
    for method in (P.METHOD_TWO_PASS, P.METHOD_LEARNED if P.learned_available() else P.METHOD_UPPER_BOUND):
        O = P.tc_fksCuckoo(P.buildHashTable(A, ctx, grouping=P.GROUP_SORT),
                           P.buildHashTable(B, ctx, grouping=P.GROUP_SORT), method=method)
        h, v = sort_coo(O.hedges_array, O.val_array)
        oh, ov, _ = orc.tc_fksCuckoo(A, B)
        oh, ov = sort_coo(oh, ov)
        assert np.array_equal(h, oh)
        assert max_rel(v, ov) <= TOL


def test_matricize_sort_uses_sorting_and_matches(ctx):
    case = golden_cases()["two_perm"]
    from tests.helpers import meta_tensors

    A, B = meta_tensors(case)
    O = P.matricize_sort_spGEMM(A, B, ctx)
    h, v = sort_coo(O.hedges_array, O.val_array)
    assert np.array_equal(h, case["coords"])
    assert max_rel(v, case["vals_dense"]) <= TOL
