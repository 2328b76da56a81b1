"""GPU engine vs oracle with a pre-drawn uniform stream (usap_set_uniforms):
both read the same buffer, so proposals, accepted moves and partitions are
equal bit for bit, and both take the same number of draws.  Reference draw
sites: graph_parallel.hpp:805-862 (merge), :1022-1046 (nodal)."""
import numpy as np
import pytest

from conftest import random_graph

pytestmark = pytest.mark.gpu


def _graph(usap, name):
    if name == "c1":
        return usap.config_graph("c1")
    return random_graph(700, 9000, seed=5, weighted=True, self_loops=True, hubs=2, isolated=7)


def _start(usap, g, init):
    return usap.scc_init(g) if init == "scc" else (np.arange(g.N, dtype=np.int32), g.N)


def _pair(usap, orc, g, lab, B, U, **params):
    eng = usap.Engine(g, **params)
    eng.set_labels(lab, B)
    eng.set_uniforms(U)
    o = orc.Oracle(g, threads=8, **params)
    o.set_labels(lab, B)
    o.set_uniforms(U)
    return eng, o


def _same_state(eng, o):
    assert eng.B == o.B
    np.testing.assert_array_equal(eng.labels(), o.labels())
    if eng.B <= 3000:
        np.testing.assert_array_equal(eng.M_dense(), o.M_dense())
    assert eng.entropy()[0] == o.entropy()[0]
    assert eng.uniforms_used() == o.uniforms_used()


@pytest.mark.parametrize("gname", ["c1", "weighted"])
@pytest.mark.parametrize("init", ["scc", "identity"])
def test_merge_rounds(usap, orc, gname, init):
    g = _graph(usap, gname)
    lab, B = _start(usap, g, init)
    U = np.random.default_rng(11).random(4 * 2 * B * 10)
    eng, o = _pair(usap, orc, g, lab, B, U, seed=3)
    for rnd in range(2):
        s, d = eng.merge_propose(1000 + rnd)   # stream arguments are ignored
        os_, od = o.block_merge(7 + rnd)
        np.testing.assert_array_equal(s, os_)
        assert d == od
        assert eng.uniforms_used() == o.uniforms_used()
        eng.merge_apply(eng.B // 2)
        assert o.carry_out(o.B // 2) == 0
        _same_state(eng, o)


@pytest.mark.parametrize("gname", ["c1", "weighted"])
@pytest.mark.parametrize("init", ["scc", "identity"])
@pytest.mark.parametrize("mode", [0, 1])
def test_nodal_batches(usap, orc, gname, init, mode):
    """Batches of uneven size, the first not at node 0: per-node proposals,
    dS and accept decisions, then the state after each batch."""
    g = _graph(usap, gname)
    lab, B = _start(usap, g, init)
    U = np.random.default_rng(12).random(4 * 2 * g.N)
    eng, o = _pair(usap, orc, g, lab, B, U, seed=4, ds_mode=mode)
    N = g.N
    for beg, end in ((137, 500), (0, 137), (500, N), (3, 300)):
        r = eng.nodal_batch(beg, end, 99)
        orr = o.nodal_batch(beg, end, 5)
        np.testing.assert_array_equal(r["s"], orr["s"])
        np.testing.assert_array_equal(r["accept"], orr["accept"])
        np.testing.assert_array_equal(r["inf"], orr["inf"])
        assert r["dS"] == orr["dS"]
        assert (r["sum"], r["sum_inf"], r["moves"]) == (orr["sum"], orr["sum_inf"], orr["moves"])
        _same_state(eng, o)


def test_sweep_equals_batches(usap, orc):
    """A device sweep takes node n's draw at pos + n, as its batches called
    one by one do."""
    g = _graph(usap, "c1")
    lab, B = _start(usap, g, "identity")
    U = np.random.default_rng(13).random(4 * 3 * g.N)
    eng, o = _pair(usap, orc, g, lab, B, U, seed=5)
    batch = 96
    for sweep in range(2):
        S = eng.entropy()[1]
        r = eng.nodal_sweep(batch, 77, S, 0.0, 0.0, 0)
        for beg in range(0, g.N, batch):
            o.nodal_batch(beg, min(beg + batch, g.N), 0)
        assert r["batches"] == -(-g.N // batch) and r["proposals"] == g.N
        _same_state(eng, o)
        assert eng.uniforms_used() == (sweep + 1) * g.N


@pytest.mark.parametrize("case", ["c1_scc", "chunked", "chunked_graphs"])
def test_partition(usap, orc, case):
    """Whole partitions: c1 from the reference's default start, and batches
    of 40 on N = 3000 from the identity start with tight slabs (chunked
    sweeps, mid-chunk rebuilds), with and without captured sweeps."""
    if case == "c1_scc":
        g = usap.config_graph("c1")
        params, eparams = dict(seed=1), {}
        lab0, B0 = usap.scc_init(g)
    else:
        g = random_graph(3000, 24000, seed=31, hubs=2)
        params = dict(seed=5, init=1, nodal_batch=40)
        eparams = dict(slab_slack=1, graphs=1 if case == "chunked_graphs" else 0)
        lab0, B0 = np.arange(g.N, dtype=np.int32), g.N
    U = np.random.default_rng(14).random(4 * 1_500_000)
    o = orc.Oracle(g, threads=16, **params)
    o.set_labels(lab0, B0)
    o.set_uniforms(U)
    B, olab, ost = o.partition()
    assert ost["error"] == 0, ost
    eng = usap.Engine(g, **params, **eparams)
    eng.set_uniforms(U)
    lab, st = eng.partition()
    assert st["error"] == 0
    assert st["B"] == B and st["S"] == ost["S"]
    np.testing.assert_array_equal(lab, olab)
    assert (st["sweeps"], st["batches"], st["nodal_proposals"], st["merge_proposals"]) == \
        (ost["sweeps"], ost["batches"], ost["nodal_proposals"], ost["merge_proposals"])
    assert eng.uniforms_used() == o.uniforms_used() == \
        st["merge_proposals"] + st["nodal_proposals"]


def test_zero_stream_closed_form(usap):
    """u = 0: every merge proposal of a block with edges is block 0 (block 1
    for block 0), of a block without edges block 0 -- derived from
    graph_parallel.hpp:805-821 without the oracle."""
    g = random_graph(400, 3000, seed=3, isolated=5)
    lab, B = usap.scc_init(g)
    eng = usap.Engine(g, K=4)
    eng.set_labels(lab, B)
    M = eng.M_dense()
    deg = M.sum(0) + M.sum(1)
    eng.set_uniforms(np.zeros(4 * B * 4))
    s, _ = eng.merge_propose(0)
    want = np.where(deg > 0, np.where(np.arange(B) == 0, 1, 0), 0)
    np.testing.assert_array_equal(s, want)
    assert eng.uniforms_used() == B * 4


def test_exhausted_stream_and_back_to_philox(usap):
    g = usap.config_graph("c1")
    lab, B = usap.scc_init(g)
    eng = usap.Engine(g, seed=1)
    eng.set_labels(lab, B)
    eng.set_uniforms(np.zeros(4 * (B * 10 - 1)))
    with pytest.raises(Exception, match="exhausted"):
        eng.merge_propose(0)
    assert eng.uniforms_used() == 0
    with pytest.raises(Exception, match="exhausted"):
        eng.nodal_sweep(250, 0, 0.0, 0.0, 0.0, 0)   # needs N draws
    eng.set_uniforms(np.random.default_rng(3).random(4 * 20000))  # a partition takes ~51k
    with pytest.raises(Exception, match="exhausted"):
        eng.partition()
    for bad in (np.zeros(6), np.array([0.5, 1.0, 0.0, 0.0])):
        with pytest.raises(Exception):
            eng.set_uniforms(bad)
    eng.set_uniforms(None)
    lab1, st1 = eng.partition()
    eng2 = usap.Engine(g, seed=1)
    lab2, st2 = eng2.partition()
    np.testing.assert_array_equal(lab1, lab2)
    assert st1["S"] == st2["S"]


def test_new_buffer_recaptures_sweep(usap):
    """A captured sweep holds the stream's device pointer: a new buffer makes
    the next sweep capture again and read the new draws."""
    g = random_graph(3000, 24000, seed=31, hubs=2)
    lab, B = np.arange(g.N, dtype=np.int32), g.N
    rng = np.random.default_rng(15)
    U1, U2 = rng.random(4 * g.N), rng.random(4 * g.N)
    eng = usap.Engine(g, seed=2, graphs=1)
    eng.set_labels(lab, B)
    eng.set_uniforms(U1)
    eng.nodal_sweep(40, 0, eng.entropy()[1], 0.0, 0.0, 0)
    eng.set_labels(lab, B)
    eng.set_uniforms(U2)
    eng.nodal_sweep(40, 0, eng.entropy()[1], 0.0, 0.0, 0)
    ref = usap.Engine(g, seed=2, graphs=0)
    ref.set_labels(lab, B)
    ref.set_uniforms(U2)
    ref.nodal_sweep(40, 0, ref.entropy()[1], 0.0, 0.0, 0)
    np.testing.assert_array_equal(eng.labels(), ref.labels())
    assert eng.entropy()[0] == ref.entropy()[0]


def test_partition_c2(usap, orc):
    """C2 (N = 20000) from the reference's default start on one shared
    pre-drawn stream (about 1.8 M draws): labels, B, S and counts."""
    g = usap.config_graph("c2")
    params = dict(seed=7)
    U = np.random.default_rng(16).random(4 * 3_000_000)
    o = orc.Oracle(g, threads=16, **params)
    o.set_labels(*o.scc())
    o.set_uniforms(U)
    B, olab, ost = o.partition()
    assert ost["error"] == 0, ost
    eng = usap.Engine(g, **params)
    eng.set_uniforms(U)
    lab, st = eng.partition()
    assert st["B"] == B and st["S"] == ost["S"]
    np.testing.assert_array_equal(lab, olab)
    assert (st["sweeps"], st["batches"], st["nodal_proposals"]) == \
        (ost["sweeps"], ost["batches"], ost["nodal_proposals"])
    assert eng.uniforms_used() == o.uniforms_used() == \
        st["merge_proposals"] + st["nodal_proposals"]
