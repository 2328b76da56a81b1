"""GPU engine vs CPU oracle, through the C-ABI (include/usap.h).

Bit-exact: labels, M, block degrees, proposed blocks, accept decisions, and
every dS / MDL as its exact int128 fixed-point value (so also within the
1e-9 relative bound on the doubles).
"""
import numpy as np
import pytest

from conftest import random_graph

pytestmark = pytest.mark.gpu


def _engine_and_oracle(usap, orc, g, labels, B, **params):
    eng = usap.Engine(g, **params)
    eng.set_labels(labels, B)
    o = orc.Oracle(g, threads=8, **params)
    o.set_labels(labels, B)
    return eng, o


def _assert_same_state(eng, o):
    assert eng.B == o.B
    np.testing.assert_array_equal(eng.labels(), o.labels())
    do, di = eng.degrees()
    odo, odi = o.degrees()
    np.testing.assert_array_equal(do, odo)
    np.testing.assert_array_equal(di, odi)
    if eng.B <= 3000:
        np.testing.assert_array_equal(eng.M_dense(), o.M_dense())
    assert eng.entropy()[0] == o.entropy()[0]


def _graphs(usap):
    return {
        "c1": usap.config_graph("c1"),
        "weighted_selfloop_hubs": random_graph(700, 9000, seed=5, weighted=True, self_loops=True,
                                               hubs=2, isolated=7),
    }


@pytest.fixture(scope="module")
def graphs(usap):
    return _graphs(usap)


@pytest.mark.parametrize("gname", ["c1", "weighted_selfloop_hubs"])
@pytest.mark.parametrize("init", ["scc", "identity"])
def test_build_and_entropy(usap, orc, graphs, gname, init):
    g = graphs[gname]
    lab, B = usap.scc_init(g) if init == "scc" else (np.arange(g.N, dtype=np.int32), g.N)
    eng, o = _engine_and_oracle(usap, orc, g, lab, B)
    _assert_same_state(eng, o)
    fx, S = eng.entropy()
    ofx, oS = o.entropy()
    assert fx == ofx and S == oS


@pytest.mark.parametrize("gname", ["c1", "weighted_selfloop_hubs"])
@pytest.mark.parametrize("init", ["scc", "identity"])
def test_merge_propose_and_apply(usap, orc, graphs, gname, init):
    g = graphs[gname]
    lab, B = usap.scc_init(g) if init == "scc" else (np.arange(g.N, dtype=np.int32), g.N)
    eng, o = _engine_and_oracle(usap, orc, g, lab, B, seed=11)
    for stream in range(2):
        s, d = eng.merge_propose(stream)
        os_, od = o.block_merge(stream)
        np.testing.assert_array_equal(s, os_)
        assert d == od
        btm = eng.B // 2
        eng.merge_apply(btm)
        assert o.carry_out(btm) == 0
        _assert_same_state(eng, o)


@pytest.mark.parametrize("K", [1, 3])
def test_merge_with_few_proposals(usap, orc, graphs, K):
    """K below the workgroup's four wavefronts: k_merge runs K wavefronts,
    which order row r's keys in LDS for the whole workgroup."""
    g = graphs["c1"]
    lab, B = usap.scc_init(g)
    eng, o = _engine_and_oracle(usap, orc, g, lab, B, seed=13, K=K)
    for stream in range(2):
        s, d = eng.merge_propose(stream)
        os_, od = o.block_merge(stream)
        np.testing.assert_array_equal(s, os_)
        assert d == od
        btm = eng.B // 2
        eng.merge_apply(btm)
        assert o.carry_out(btm) == 0
        _assert_same_state(eng, o)


@pytest.mark.parametrize("gname", ["c1", "weighted_selfloop_hubs"])
@pytest.mark.parametrize("init", ["scc", "identity"])
@pytest.mark.parametrize("mode", [0, 1])
def test_nodal_batches(usap, orc, graphs, gname, init, mode):
    g = graphs[gname]
    lab, B = usap.scc_init(g) if init == "scc" else (np.arange(g.N, dtype=np.int32), g.N)
    eng, o = _engine_and_oracle(usap, orc, g, lab, B, seed=7, ds_mode=mode)
    if init == "identity":  # one merge round first, so blocks hold several nodes
        eng.merge_propose(0)
        o.block_merge(0)
        eng.merge_apply(eng.B // 2)
        o.carry_out(o.B // 2)
    batch = g.N // 4
    stream = 0
    for sweep in range(2):
        for beg in range(0, g.N, batch):
            end = min(beg + batch, g.N)
            r = eng.nodal_batch(beg, end, stream)
            ro = o.nodal_batch(beg, end, stream)
            np.testing.assert_array_equal(r["s"], ro["s"])
            np.testing.assert_array_equal(r["accept"], ro["accept"])
            np.testing.assert_array_equal(r["inf"], np.where(ro["inf"] == -1, -1, 0))
            fin = ro["inf"] == 0
            assert [d for d, f in zip(r["dS"], fin) if f] == [d for d, f in zip(ro["dS"], fin) if f]
            assert r["moves"] == ro["moves"]
            assert r["sum_inf"] == ro["sum_inf"]
            if not ro["sum_inf"]:
                assert r["sum"] == ro["sum"]
            _assert_same_state(eng, o)
            stream += 1


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("init", [0, 1])
def test_partition_c1(usap, orc, graphs, mode, init):
    g = graphs["c1"]
    params = dict(seed=3, ds_mode=mode, init=init)
    eng = usap.Engine(g, **params)
    lab, st = eng.partition()
    o = orc.Oracle(g, threads=8, **params)
    if init == 0:
        l0, b0 = o.scc()
    else:
        l0, b0 = np.arange(g.N, dtype=np.int32), g.N
    o.set_labels(l0, b0)
    B, olab, ost = o.partition()
    assert st["error"] == 0 and ost["error"] == 0
    assert st["B"] == B
    np.testing.assert_array_equal(lab, olab)
    assert st["S"] == ost["S"]
    assert st["sweeps"] == ost["sweeps"] and st["batches"] == ost["batches"]
    assert st["accepted_moves"] == ost["moves"]


def test_partition_weighted_selfloops(usap, orc, graphs):
    g = graphs["weighted_selfloop_hubs"]
    params = dict(seed=9, ds_mode=1, init=0)
    eng = usap.Engine(g, **params)
    lab, st = eng.partition()
    o = orc.Oracle(g, threads=8, **params)
    l0, b0 = o.scc()
    o.set_labels(l0, b0)
    B, olab, ost = o.partition()
    assert st["error"] == 0 and ost["error"] == 0
    assert st["B"] == B
    np.testing.assert_array_equal(lab, olab)
    assert st["S"] == ost["S"]
