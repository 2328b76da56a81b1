"""CPU tests of the product's host side: the C-ABI library loads and exports
every declared entry point, graph ingest (TSV, CSR/CSC order), the SCC
initial partition, the DCSBM generator and the quality metrics."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, random_graph


def test_abi_exports_every_declared_symbol(usap):
    from usap_amd import _lib
    header = open(os.path.join(ROOT, "include", "usap.h")).read()
    declared = set(re.findall(r"\b(usap_[A-Za-z0-9_]+)\s*\(", header))
    assert len(declared) >= 25
    L = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert declared == bound


def test_engine_without_device_fails_loudly(usap):
    g = usap.config_graph("c1")
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError):
        usap.Engine(g)


def test_csr_keeps_file_order(usap):
    src = np.array([2, 0, 2, 1, 0, 2], np.int32)
    dst = np.array([0, 2, 1, 2, 1, 0], np.int32)
    w = np.array([1, 2, 3, 4, 5, 6], np.int32)
    g = usap.Graph.from_edges(src, dst, 3, w)
    np.testing.assert_array_equal(g.out_ptr, [0, 2, 3, 6])
    np.testing.assert_array_equal(g.out_idx, [2, 1, 2, 0, 1, 0])
    np.testing.assert_array_equal(g.out_w, [2, 5, 4, 1, 3, 6])
    np.testing.assert_array_equal(g.in_ptr, [0, 2, 4, 6])
    np.testing.assert_array_equal(g.in_idx, [2, 2, 2, 0, 0, 1])
    np.testing.assert_array_equal(g.in_w, [1, 6, 3, 5, 2, 4])


def test_tsv_loader(usap, tmp_path):
    p = tmp_path / "g.tsv"
    p.write_text("1\t3\t1\n2\t1\t2\n3\t2\t1\n3\t4\t1\n")
    (tmp_path / "g_truePartition.tsv").write_text("1\t1\n2\t1\n3\t2\n4\t2\n")
    g = usap.load_tsv(str(p), str(tmp_path / "g_truePartition.tsv"))
    assert g.N == 4 and g.E == 4  # N counts the sink-only node 4
    np.testing.assert_array_equal(g.out_idx, [2, 0, 1, 3])
    np.testing.assert_array_equal(g.out_w, [1, 2, 1, 1])
    np.testing.assert_array_equal(g.truth, [0, 0, 1, 1])
    gp = usap.Graph_P(str(tmp_path / "g"))
    assert gp.num_nodes() == 4 and gp.num_edges() == 4
    np.testing.assert_array_equal(gp.truePartitions, [0, 0, 1, 1])
    assert gp.num_agg_proposals_per_block == 10 and gp.dfs_depth == 20


def test_tsv_loader_rejects_garbage(usap, tmp_path):
    p = tmp_path / "bad.tsv"
    p.write_text("1\tx\t1\n")
    with pytest.raises(RuntimeError):
        usap.load_tsv(str(p))
    with pytest.raises(RuntimeError):
        usap.load_tsv(str(tmp_path / "missing.tsv"))


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_generator(usap, name):
    cfg = usap.CONFIGS[name]
    g = usap.config_graph(name)
    g2 = usap.config_graph(name)
    np.testing.assert_array_equal(g.out_idx, g2.out_idx)  # deterministic
    np.testing.assert_array_equal(g.truth, g2.truth)
    assert g.N == cfg["N"]
    sizes = np.bincount(g.truth)
    assert len(sizes) == cfg["K"] and sizes.sum() == g.N and sizes.min() >= cfg["min_block"]
    src, dst = g.edges()
    assert not np.any(src == dst)
    key = src.astype(np.int64) * g.N + dst
    assert len(np.unique(key)) == len(key)
    deg = np.diff(g.out_ptr)
    assert deg.min() >= cfg["dmin"] and deg.max() <= cfg["dmax"]
    within = np.mean(g.truth[src] == g.truth[dst])
    assert abs(within - cfg["p_in"]) < 0.03
    # power-law mean out-degree on [10, 100] with exponent 2.5 is about 21.2
    assert 19 < deg.mean() < 23.5
    other = usap.config_graph(name, seed=cfg["seed"] + 100)
    assert not np.array_equal(other.out_idx[:100], g.out_idx[:100])


def test_metrics_against_sklearn(usap):
    from sklearn.metrics import adjusted_rand_score, normalized_mutual_info_score
    rng = np.random.default_rng(0)
    for _ in range(10):
        t = rng.integers(0, 7, 300)
        p = np.where(rng.random(300) < 0.7, t, rng.integers(0, 9, 300))
        assert abs(usap.nmi(t, p) - normalized_mutual_info_score(t, p)) < 1e-12
        prec, rec, ri, ari = usap.pairwise_precision_recall(t, p)
        assert abs(ari - adjusted_rand_score(t, p)) < 1e-12
        assert 0 <= prec <= 1 and 0 <= rec <= 1
    t = np.array([0, 0, 1, 1])
    assert usap.pairwise_precision_recall(t, t)[:2] == (1.0, 1.0)
    prec, rec, _, _ = usap.pairwise_precision_recall(t, np.array([0, 0, 0, 0]))
    assert rec == 1.0 and abs(prec - 2 / 6) < 1e-12


def _by_source(g):
    """The same edges listed by source (the file order of a sorted edge list):
    every in-list then names its sources in ascending order."""
    import usap_amd
    src = np.repeat(np.arange(g.N, dtype=np.int32), np.diff(g.out_ptr))
    return usap_amd.Graph.from_edges(src, g.out_idx, g.N)


@pytest.mark.parametrize("kind", ["shuffled", "by_source", "generator", "chain"])
@pytest.mark.parametrize("depth", [1, 20, 300])
def test_scc_init_matches_oracle(usap, orc, kind, depth):
    """usap_scc_init (a cursor per stacked node instead of one stack entry per
    edge; the in-lists taken as the transpose when they are source-ascending,
    else a counting sort) against the oracle's literal restatement of
    _scc/_dfs_itr (graph_parallel.hpp:514-603): same labels and B, with
    multi-edges, self-loops, hubs, isolated nodes and a DFS as deep as N."""
    if kind == "shuffled":
        g = random_graph(2500, 15000, seed=9, self_loops=True, hubs=2, isolated=6)
    elif kind == "by_source":
        g = _by_source(random_graph(2500, 15000, seed=9, self_loops=True, hubs=2, isolated=6))
    elif kind == "generator":
        g = usap.config_graph("c1")
    else:  # one long path with a few back edges: the first DFS goes N deep
        n = 30000
        src = np.concatenate([np.arange(n - 1), [n - 1, 17000, 900]])
        dst = np.concatenate([np.arange(1, n), [0, 5, 29000]])
        g = usap.Graph.from_edges(src, dst, n)
    lab, B = usap.scc_init(g, depth)
    olab, oB = orc.Oracle(g).scc(depth)
    assert B == oB
    np.testing.assert_array_equal(lab, olab)


def test_scc_init_without_in_lists(usap, orc):
    """The in-lists are optional for usap_scc_init: without them the transpose
    is built from the out-lists, with the same result."""
    from usap_amd import _lib
    g = _by_source(random_graph(1500, 9000, seed=4, self_loops=True, hubs=1))
    gs = g.as_struct()
    gs.in_ptr = None
    gs.in_idx = None
    lab = np.zeros(g.N, np.int32)
    B = C.c_int32()
    assert _lib.lib().usap_scc_init(C.byref(gs), 20, lab, C.byref(B)) == 0
    olab, oB = orc.Oracle(g).scc(20)
    assert B.value == oB
    np.testing.assert_array_equal(lab, olab)
