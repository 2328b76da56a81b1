"""The device-side sweep (usap_nodal_sweep: batches plus the reference's
update step and convergence test, graph_parallel.hpp:397-468, on the GPU)
against the oracle driven batch by batch with that step on the host."""
import math

import numpy as np
import pytest

from conftest import random_graph

pytestmark = pytest.mark.gpu


def fx_to_double(v):
    """The engine's int128 -> double: |v| split in 64-bit halves, two ldexp, one add."""
    neg = v < 0
    m = -v if neg else v
    d = math.ldexp(float(m >> 64), 16) + math.ldexp(float(m & ((1 << 64) - 1)), -48)
    return -d if neg else d


def oracle_sweep(o, N, batch, stream0, S, mean, thr, check):
    cum, inf, done, moves = 0, False, 0, 0
    dS = 0.0
    converged = False
    for beg in range(0, N, batch):
        r = o.nodal_batch(beg, min(beg + batch, N), stream0 + done)
        done += 1
        moves += r["moves"]
        if r["sum_inf"]:
            inf = True
        else:
            cum += r["sum"]
        dS = -math.inf if inf else fx_to_double(cum)
        S += dS
        if check and -mean < thr * S:
            converged = True
            break
    return dict(batches=done, moves=moves, S=S, dS_sum=dS, converged=converged, cum=cum)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("slack", [4, 1])
def test_sweep_equals_batch_loop(usap, orc, mode, slack):
    g = usap.config_graph("c1")
    eng = usap.Engine(g, seed=4, ds_mode=mode, slab_slack=slack)
    o = orc.Oracle(g, threads=8, seed=4, ds_mode=mode)
    lab = np.arange(g.N, dtype=np.int32)
    eng.set_labels(lab, g.N)
    o.set_labels(lab, g.N)
    for e in (eng, o):  # one merge round: blocks of ~2 nodes, hashed rows
        if e is eng:
            e.merge_propose(0)
            e.merge_apply(g.N // 2)
        else:
            e.block_merge(0)
            e.carry_out(g.N // 2)
    stream = 0
    for itr, (mean, check) in enumerate([(0.0, 0), (-50.0, 1), (-1e9, 1), (-1.0, 1)]):
        _, S0 = o.entropy()
        a = eng.nodal_sweep(50, stream, S0, mean, 1e-4, check)
        b = oracle_sweep(o, g.N, 50, stream, S0, mean, 1e-4, check)
        assert a["batches"] == b["batches"] and bool(a["converged"]) == b["converged"]
        assert a["moves"] == b["moves"]
        assert a["S"] == b["S"] and (a["dS_sum"] == b["dS_sum"] or math.isinf(b["dS_sum"]))
        if not math.isinf(b["dS_sum"]):
            assert a["dS_fx"] == b["cum"]
        np.testing.assert_array_equal(eng.labels(), o.labels())
        np.testing.assert_array_equal(eng.M_dense(), o.M_dense())
        stream += a["batches"]


@pytest.mark.parametrize("gname", ["sparse", "weighted"])
def test_partition_with_tight_slabs(usap, orc, gname):
    """slab_slack=1 makes hashed rows fill up often: rows are moved to larger
    slabs between batches (and M is rebuilt when the pool runs out); the
    partition must not change.  The sparse graph keeps B large against the
    rows' key counts (a row is direct once 8 x its slab reaches B), so its
    rows stay hashed."""
    g = random_graph(3000, 24000, seed=6, hubs=2) if gname == "sparse" else random_graph(
        3000, 24000, seed=5, weighted=True, self_loops=True, hubs=2, isolated=7)
    params = dict(seed=2, init=1, ds_mode=1)
    eng = usap.Engine(g, slab_slack=1, **params)
    lab, st = eng.partition()
    assert st["grows"] > 0  # rows were grown
    o = orc.Oracle(g, threads=8, **params)
    o.set_labels(np.arange(g.N, dtype=np.int32), g.N)
    B, olab, ost = o.partition()
    assert st["B"] == B and st["S"] == ost["S"]
    np.testing.assert_array_equal(lab, olab)


@pytest.mark.parametrize("gname,slack", [("c1", 4), ("weighted", 1)])
def test_graph_replay_equals_direct_launches(usap, gname, slack):
    """Sweeps replayed from a captured HIP graph (the default) and launched
    kernel by kernel give the same partition; the captured sweep is replayed
    many times with host work (entropy, merges, snapshots) in between, and
    re-captured whenever B or the M pools change."""
    g = usap.config_graph("c1") if gname == "c1" else random_graph(
        700, 9000, seed=5, weighted=True, self_loops=True, hubs=2, isolated=7)
    res = []
    for graphs in (1, 0):
        eng = usap.Engine(g, seed=6, init=1, ds_mode=1, slab_slack=slack, graphs=graphs)
        lab, st = eng.partition()
        eng.close()
        res.append((lab, st))
    np.testing.assert_array_equal(res[0][0], res[1][0])
    for k in ("S", "B", "sweeps", "batches", "accepted_moves"):
        assert res[0][1][k] == res[1][1][k]
