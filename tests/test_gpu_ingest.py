"""Graph ingest on the device (usap_create_from_edges, usap_append_edges)
against the host CSR/CSC build (usap_build_csr, graph_parallel.hpp:251-317
neighbour order :290-294): an engine built from an edge list, or grown by
increments, holds the same graph as one created from the host CSR/CSC of the
same list -- M, degrees, MDL, and every later proposal (partitions compared
bit for bit)."""
import numpy as np
import pytest

import usap_amd
from usap_amd._lib import USAP_EINVAL, UsapError

pytestmark = pytest.mark.gpu


def _edges(N, E, seed, weighted=False, self_loops=True, hubs=1):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, N, E)
    dst = rng.integers(0, N, E)
    if not self_loops:
        keep = src != dst
        src, dst = src[keep], dst[keep]
    for h in range(hubs):  # a node beyond the LDS table, in both directions
        k = 500
        o = rng.integers(0, N, k)
        src = np.concatenate([src, np.full(k, h), o])
        dst = np.concatenate([dst, o, np.full(k, h)])
    w = rng.integers(1, 4, len(src)).astype(np.int32) if weighted else None
    return src.astype(np.int32), dst.astype(np.int32), w


def _same_state(a, b):
    assert a.B == b.B
    np.testing.assert_array_equal(a.labels(), b.labels())
    for x, y in zip(a.M_coo(), b.M_coo()):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(a.degrees(), b.degrees()):
        np.testing.assert_array_equal(x, y)
    assert a.entropy() == b.entropy()


@pytest.mark.parametrize("weighted", [False, True])
def test_create_from_edges_equals_host_csr(weighted):
    src, dst, w = _edges(2500, 30000, seed=41, weighted=weighted)
    g = usap_amd.Graph.from_edges(src, dst, 2500, w)
    params = dict(seed=3, init=0)
    a = usap_amd.Engine(g, **params)
    b = usap_amd.Engine(g, edges=(src, dst, w), **params)
    lab, B = usap_amd.scc_init(g)
    a.set_labels(lab, B)
    b.set_labels(lab, B)
    _same_state(a, b)
    la, sa = a.partition()
    lb, sb = b.partition()  # SCC init from the host copy refreshed off the device
    assert sa["S"] == sb["S"] and sa["B"] == sb["B"] and sa["batches"] == sb["batches"]
    np.testing.assert_array_equal(la, lb)


@pytest.mark.parametrize("weighted", [False, True])
def test_append_increments_equal_set_graph(weighted):
    """Ten increments of a shuffled edge list: after each, the grown engine
    equals usap_set_graph with the CSR/CSC of the whole prefix, and so do a
    resumed nodal phase and, at the end, a full partition."""
    N = 3000
    src, dst, w = _edges(N, 36000, seed=42, weighted=weighted)
    perm = np.random.default_rng(5).permutation(len(src))
    src, dst = src[perm], dst[perm]
    w = None if w is None else w[perm]
    cuts = [int(np.ceil(len(src) * k / 10)) for k in range(11)]
    params = dict(seed=4, init=0)
    g1 = usap_amd.Graph.from_edges(src[:cuts[1]], dst[:cuts[1]], N,
                                   None if w is None else w[:cuts[1]])
    grown = usap_amd.Engine(g1, edges=(src[:cuts[1]], dst[:cuts[1]],
                                       None if w is None else w[:cuts[1]]), **params)
    ref = usap_amd.Engine(g1, **params)
    grown.partition()
    ref.partition()
    for k in range(2, 11):
        a, b = cuts[k - 1], cuts[k]
        print(f"increment {k}: edges [{a},{b})", flush=True)
        grown.append_edges(src[a:b], dst[a:b], None if w is None else w[a:b])
        gk = usap_amd.Graph.from_edges(src[:b], dst[:b], N, None if w is None else w[:b])
        ref.set_graph(gk)
        _same_state(grown, ref)
        la, sa = grown.resume(k << 40)
        lb, sb = ref.resume(k << 40)
        assert sa["S"] == sb["S"] and sa["batches"] == sb["batches"]
        np.testing.assert_array_equal(la, lb)
    la, sa = grown.partition()
    lb, sb = ref.partition()
    assert sa["S"] == sb["S"] and sa["B"] == sb["B"]
    np.testing.assert_array_equal(la, lb)


def test_append_edge_cases():
    src, dst, _ = _edges(500, 4000, seed=43, hubs=0)
    g = usap_amd.Graph.from_edges(src, dst, 500)
    e = usap_amd.Engine(g, seed=1)
    lab, B = usap_amd.scc_init(g)
    e.set_labels(lab, B)
    before = e.entropy()
    e.append_edges(np.zeros(0, np.int32), np.zeros(0, np.int32))  # nothing to add
    assert e.entropy() == before
    with pytest.raises(UsapError) as ex:
        e.append_edges(np.array([0], np.int32), np.array([500], np.int32))
    assert ex.value.code == USAP_EINVAL
    with pytest.raises(UsapError) as ex:  # weights for an unweighted graph
        e.append_edges(np.array([0], np.int32), np.array([1], np.int32), np.array([2], np.int32))
    assert ex.value.code == USAP_EINVAL
    assert e.entropy() == before  # rejected appends leave the graph alone
    # an engine from no edges at all, then everything in one increment
    z = usap_amd.Engine(g, edges=(np.zeros(0, np.int32), np.zeros(0, np.int32), None), seed=1)
    z.append_edges(src, dst)
    z.set_labels(lab, B)
    assert z.entropy() == before


def test_append_c2_ten_increments(usap):
    """The N=20000 workload's edges, shuffled, in ten increments (as the
    served stream delivers them): the grown engine equals the whole graph's."""
    g = usap.config_graph("c2")
    src, dst = g.edges()
    perm = np.random.default_rng(2).permutation(len(src))
    src, dst = src[perm], dst[perm]
    cuts = [int(np.ceil(len(src) * k / 10)) for k in range(11)]
    eng = usap.Engine(g, edges=(src[:cuts[1]], dst[:cuts[1]], None), seed=2, init=0)
    for k in range(2, 11):
        eng.append_edges(src[cuts[k - 1]:cuts[k]], dst[cuts[k - 1]:cuts[k]])
    full = usap.Engine(usap.Graph.from_edges(src, dst, g.N), seed=2, init=0)
    la, sa = eng.partition()
    lb, sb = full.partition()
    assert sa["S"] == sb["S"] and sa["B"] == sb["B"]
    np.testing.assert_array_equal(la, lb)
