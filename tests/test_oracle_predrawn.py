"""The oracle's pre-drawn uniform stream (orc_set_uniforms), pinned without
the oracle's own proposal code.

Constant streams make every proposal a closed form of the dense M:
  u = 0            every index is 0 and the uniform branch is always taken;
  u = 1 - 2^-53    every index is W - 1 and the uniform branch is never taken
                   for a block with edges, so a multinomial pick lands on the
                   largest block with non-zero weight.
The draw-to-proposal mapping (block r, proposal j -> draw pos + r*K + j;
node n of batch [beg, end) -> draw pos + n - beg) is pinned by changing one
draw and watching which proposal changes.  Reference draw sites:
graph_parallel.hpp:805-862 (merge), :1022-1046 (nodal).
"""
import numpy as np
import pytest

from conftest import random_graph

UMAX = 1.0 - 2.0 ** -53
K = 4


def _graphs(usap):
    return {
        "plain": random_graph(400, 3000, seed=3, isolated=5),
        "weighted": random_graph(300, 2500, seed=4, weighted=True, self_loops=True, isolated=3),
    }


@pytest.fixture(scope="module")
def graphs(usap):
    return _graphs(usap)


def _oracle(orc, g, init, usap):
    o = orc.Oracle(g, threads=4, K=K, seed=9)
    if init == "scc":
        o.set_labels(*o.scc())
    else:
        o.set_labels(np.arange(g.N, dtype=np.int32), g.N)
    return o


def _merge_max(M, r):
    """Merge proposal of block r under u = 1 - 2^-53 (graph_parallel.hpp:805-862)."""
    B = len(M)
    w = M[r] + M[:, r]
    if w.sum() == 0:
        return B - 1                       # uniform block: index B - 1
    u = np.nonzero(w)[0].max()             # neighbour block: the last unit
    wu = M[u] + M[:, u]
    wu[r] = 0                              # the multinomial excludes r
    if wu.sum() == 0:                      # no weight left: uniform block other than r
        s = B - 1 + (r == B - 1)
        return 0 if s == B else s
    return np.nonzero(wu)[0].max()


def _nodal_max(g, lab, M, n):
    """Nodal proposal of node n under u = 1 - 2^-53 (graph_parallel.hpp:1010-1046):
    the last unit of the list (in-neighbours after out-neighbours), then the
    largest block with weight in that block's row and column of M."""
    B = len(M)
    o0, o1 = g.out_ptr[n], g.out_ptr[n + 1]
    i0, i1 = g.in_ptr[n], g.in_ptr[n + 1]
    if o1 == o0 and i1 == i0:
        return B - 1
    v = g.in_idx[i1 - 1] if i1 > i0 else g.out_idx[o1 - 1]
    u = lab[v]
    wu = M[u] + M[:, u]
    return np.nonzero(wu)[0].max()


@pytest.mark.parametrize("gname", ["plain", "weighted"])
@pytest.mark.parametrize("init", ["scc", "identity"])
def test_merge_constant_streams(usap, orc, graphs, gname, init):
    g = graphs[gname]
    o = _oracle(orc, g, init, usap)
    B = o.B
    M = o.M_dense()
    deg = M.sum(0) + M.sum(1)
    o.set_uniforms(np.zeros(4 * B * K))
    s, _ = o.block_merge(0)
    want = np.where(deg > 0, np.where(np.arange(B) == 0, 1, 0), 0)
    np.testing.assert_array_equal(s, want)
    assert o.uniforms_used() == B * K
    o.set_uniforms(np.full(4 * B * K, UMAX))
    s, _ = o.block_merge(0)
    np.testing.assert_array_equal(s, [_merge_max(M, r) for r in range(B)])


@pytest.mark.parametrize("gname", ["plain", "weighted"])
@pytest.mark.parametrize("init", ["scc", "identity"])
def test_nodal_constant_streams(usap, orc, graphs, gname, init):
    g = graphs[gname]
    o = _oracle(orc, g, init, usap)
    lab, B, M = o.labels(), o.B, o.M_dense()
    beg, end = 37, 211
    o.set_uniforms(np.full(4 * (end - beg), UMAX))
    r = o.nodal_batch(beg, end, 0)
    np.testing.assert_array_equal(r["s"], [_nodal_max(g, lab, M, n) for n in range(beg, end)])
    assert o.uniforms_used() == end - beg
    o.set_uniforms(np.zeros(4 * g.N))
    r = o.nodal_batch(0, g.N, 0)
    np.testing.assert_array_equal(r["s"], np.zeros(g.N, np.int32))


def test_merge_draw_mapping(usap, orc, graphs):
    """Zeroing draw pos + r*K + j changes proposal j of block r only: it then
    equals the u = 0 closed form, and every other block's best merge stays."""
    g = graphs["plain"]
    o = _oracle(orc, g, "scc", usap)
    B = o.B
    rng = np.random.default_rng(1)
    pre = 3  # a batch consumed first moves the cursor
    U = rng.random(4 * (pre + B * K))
    o.set_uniforms(U)
    o.nodal_batch(0, pre, 0)
    r, j = B // 2, K - 1
    before = [o.merge_propose_one(r, jj, 0) for jj in range(K)]  # reads, takes nothing
    s0, _ = o.block_merge(0)
    o2 = _oracle(orc, g, "scc", usap)
    U2 = U.copy()
    U2[4 * (pre + r * K + j):4 * (pre + r * K + j + 1)] = 0.0
    o2.set_uniforms(U2)
    o2.nodal_batch(0, pre, 0)
    assert o2.merge_propose_one(r, j, 0)[0] == (1 if r == 0 else 0)
    for jj in range(K - 1):
        assert o2.merge_propose_one(r, jj, 0) == before[jj]
    assert o2.uniforms_used() == pre
    s2, _ = o2.block_merge(0)
    keep = np.arange(B) != r
    np.testing.assert_array_equal(s2[keep], s0[keep])


def test_nodal_draw_mapping(usap, orc, graphs):
    """Node n of batch [beg, end) takes draw pos + n - beg: zeroing that draw
    turns its proposal into block 0 and leaves every other node's."""
    g = graphs["plain"]
    o = _oracle(orc, g, "identity", usap)
    rng = np.random.default_rng(2)
    beg, end, pos = 120, 260, 5
    U = rng.random(4 * (pos + end - beg))
    n = 200
    U2 = U.copy()
    U2[4 * (pos + n - beg):4 * (pos + n - beg + 1)] = 0.0
    out = []
    for buf in (U, U2):
        o = _oracle(orc, g, "identity", usap)
        o.set_uniforms(buf)
        o.nodal_batch(0, pos, 0)
        out.append(o.nodal_batch(beg, end, 0)["s"])
    assert out[1][n - beg] == 0
    keep = np.arange(end - beg) != n - beg
    np.testing.assert_array_equal(out[1][keep], out[0][keep])


def test_partition_consumes_its_proposals(usap, orc):
    """A partition takes one draw per merge proposal and per nodal proposal,
    and stops with error -4 when the stream is too short."""
    g = usap.config_graph("c1")
    o = orc.Oracle(g, threads=4, seed=1)
    o.set_labels(*o.scc())
    o.set_uniforms(np.random.default_rng(3).random(4 * 80000))
    B, lab, st = o.partition()
    assert st["error"] == 0 and B > 0
    assert o.uniforms_used() == st["merge_proposals"] + st["nodal_proposals"]
    o2 = orc.Oracle(g, threads=4, seed=1)
    o2.set_labels(*o2.scc())
    o2.set_uniforms(np.random.default_rng(3).random(4 * 20000))
    B2, _, st2 = o2.partition()
    assert B2 == -4 and st2["error"] == -4


def test_bad_buffers(usap, orc, graphs):
    o = _oracle(orc, graphs["plain"], "scc", usap)
    for bad in (np.zeros(6), np.array([0.1, 0.2, 1.0, 0.3]), np.array([0.1, -0.5, 0.2, 0.3]),
                np.array([0.1, np.nan, 0.2, 0.3])):
        with pytest.raises(ValueError):
            o.set_uniforms(bad)
    o.set_uniforms(np.zeros(8))
    with pytest.raises(RuntimeError):
        o.block_merge(0)   # needs B*K draws
    assert o.uniforms_used() == 0
