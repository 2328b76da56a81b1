"""GPU engine against the oracle on the configured workloads' code paths.

* N=20000 DCSBM (32 planted blocks, p_in 0.6): identity start, one merge
  round to B=10000 (the M build by per-edge inserts, B > 4096 with blocks of
  several nodes), then 4096-node batches compared proposal by proposal; full
  partitions from the SCC start and from the identity start with 4096-node
  batches, the latter phase by phase (the regime where early proposals are
  nearly uniform: B / (d_u + B) close to 1).
* Sweeps long enough to go out in launch chunks (more than 32 batches), with
  tight slabs, against the oracle; a sweep whose convergence test fires in
  the middle of a chunk.
* The N=500k DCSBM at full size: after merge rounds and nodal sweeps, the
  device's M equals the edge counts of its own labels, the block degrees are
  M's row and column sums, and usap_entropy equals the MDL recomputed on the
  host from that M.
"""
import math

import numpy as np
import pytest

from conftest import random_graph

pytestmark = pytest.mark.gpu


def _flush(*a):
    print(*a, flush=True)


def _coo_from_labels(g, lab, B):
    """M[a][b] = edges u->v with lab[u] = a, lab[v] = b, as (keys a*B+b, counts),
    keys ascending."""
    src, dst = g.edges()
    key = lab[src].astype(np.int64) * B + lab[dst].astype(np.int64)
    w = np.ones(len(key), np.int64) if g.out_w is None else g.out_w.astype(np.int64)
    order = np.argsort(key, kind="stable")
    key, w = key[order], w[order]
    if len(key) == 0:
        return key, w
    starts = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    return key[starts], np.add.reduceat(w, starts)


def _sum_L(vals, lam):
    """sum of m * lambda(m) over the values m > 0, exactly (Python ints over
    the distinct values)."""
    v = vals[vals > 0]
    u, c = np.unique(v, return_counts=True)
    return sum(int(ci) * int(ui) * int(lam[ui]) for ui, ci in zip(u, c))


def _check_state(usap, eng, g, lam):
    """Device M, degrees and MDL against the labels the device holds."""
    B = eng.B
    lab = eng.labels()
    assert lab.min() >= 0 and lab.max() < B
    rows, cols, vals = eng.M_coo()
    key = rows.astype(np.int64) * B + cols.astype(np.int64)
    want_key, want_val = _coo_from_labels(g, lab, B)
    np.testing.assert_array_equal(key, want_key)
    np.testing.assert_array_equal(vals, want_val)
    do, di = eng.degrees()
    np.testing.assert_array_equal(do, np.bincount(rows, weights=vals, minlength=B).astype(np.int64))
    np.testing.assert_array_equal(di, np.bincount(cols, weights=vals, minlength=B).astype(np.int64))
    fx, S = eng.entropy()
    want = _sum_L(do.astype(np.int64), lam) + _sum_L(di.astype(np.int64), lam) - \
        _sum_L(vals.astype(np.int64), lam)
    assert fx == want
    assert S == usap.model_S(g.N, g.E, B) + _fx_to_double(want)


def _fx_to_double(v):
    """The engine's int128 -> double: |v| split in 64-bit halves, two ldexp, one add."""
    neg = v < 0
    m = -v if neg else v
    d = math.ldexp(float(m >> 64), 16) + math.ldexp(float(m & ((1 << 64) - 1)), -48)
    return -d if neg else d


@pytest.fixture(scope="module")
def c2(usap):
    return usap.config_graph("c2")


def test_c2_large_B_build_and_batches(usap, orc, c2):
    """B = 10000 after one merge round of the identity start: the rebuild
    inserts edge by edge (B > 4096, blocks of two nodes, hashed rows); then
    eight 4096-node batches against the frozen M, every proposal, accept flag
    and dS compared, and the state after each batch."""
    g = c2
    lam = usap.lambda_table(int(g.E))
    lab0 = np.arange(g.N, dtype=np.int32)
    eng = usap.Engine(g, seed=21, init=1)
    eng.set_labels(lab0, g.N)
    o = orc.Oracle(g, threads=16, seed=21)
    o.set_labels(lab0, g.N)
    s, d = eng.merge_propose(0)
    os_, od = o.block_merge(0)
    np.testing.assert_array_equal(s, os_)
    assert d == od
    eng.merge_apply(g.N // 2)
    assert o.carry_out(g.N // 2) == 0
    assert eng.B == o.B == g.N // 2
    np.testing.assert_array_equal(eng.labels(), o.labels())
    _check_state(usap, eng, g, lam)
    stream = 0
    for sweep in range(2):
        for beg in range(0, g.N, 4096):
            end = min(beg + 4096, g.N)
            _flush(f"c2 batch sweep {sweep} [{beg},{end})")
            r = eng.nodal_batch(beg, end, stream)
            ro = o.nodal_batch(beg, end, stream)
            np.testing.assert_array_equal(r["s"], ro["s"])
            np.testing.assert_array_equal(r["accept"], ro["accept"])
            np.testing.assert_array_equal(r["inf"], np.where(ro["inf"] == -1, -1, 0))
            fin = ro["inf"] == 0
            assert [x for x, f in zip(r["dS"], fin) if f] == [x for x, f in zip(ro["dS"], fin) if f]
            assert r["moves"] == ro["moves"] and r["sum_inf"] == ro["sum_inf"]
            if not ro["sum_inf"]:
                assert r["sum"] == ro["sum"]
            np.testing.assert_array_equal(eng.labels(), o.labels())
            do, di = eng.degrees()
            odo, odi = o.degrees()
            np.testing.assert_array_equal(do, odo)
            np.testing.assert_array_equal(di, odi)
            stream += 1
        _check_state(usap, eng, g, lam)
        assert eng.entropy()[0] == o.entropy()[0]


def test_c2_partition_scc(usap, orc, c2):
    """Full partition from the reference's SCC start, batches of N/4."""
    params = dict(seed=2, init=0)
    eng = usap.Engine(c2, **params)
    lab, st = eng.partition()
    o = orc.Oracle(c2, threads=16, **params)
    o.set_labels(*o.scc())
    B, olab, ost = o.partition()
    assert st["error"] == 0 and ost["error"] == 0
    assert st["B"] == B and st["S"] == ost["S"]
    np.testing.assert_array_equal(lab, olab)
    assert (st["sweeps"], st["batches"], st["accepted_moves"]) == \
        (ost["sweeps"], ost["batches"], ost["moves"])
    assert [(p["B"], p["S"], p["sweeps"]) for p in eng.partition_log()] == \
        [(p["B"], p["S"], p["sweeps"]) for p in o.partition_log()]


def test_c2_partition_identity_4096(usap, orc, c2):
    """The identity start with 4096-node batches (as the N=500k workload):
    the first phases run at B in the thousands, where a proposal follows M
    with probability d_u / (d_u + B) only, so most are uniform.  Engine and
    oracle agree phase by phase on B, S, sweeps, batches and moves, and on the
    final labels."""
    params = dict(seed=22, init=1, nodal_batch=4096)
    eng = usap.Engine(c2, **params)
    lab, st = eng.partition()
    o = orc.Oracle(c2, threads=16, **params)
    o.set_labels(np.arange(c2.N, dtype=np.int32), c2.N)
    B, olab, ost = o.partition()
    assert st["error"] == 0 and ost["error"] == 0
    el, ol = eng.partition_log(), o.partition_log()
    assert len(el) == len(ol)
    for i, (a, b) in enumerate(zip(el, ol)):
        assert (a["B"], a["S"], a["sweeps"], a["moves"]) == \
            (b["B"], b["S"], b["sweeps"], b["moves"]), i
    assert st["B"] == B and st["S"] == ost["S"]
    np.testing.assert_array_equal(lab, olab)
    assert st["batches"] == ost["batches"]


@pytest.mark.parametrize("mode", [0, 1])
def test_chunked_sweeps_partition(usap, orc, mode):
    """Batches of 40 nodes on N = 3000 (75 a sweep, so sweeps that test
    convergence go out in chunks of 16 batches, one chunk ahead of the host)
    with slab_slack 1 (rows grow between batches, M is rebuilt and the sweep
    resumed mid-chunk when the pool runs out): the same partition as the
    oracle's, sweep and batch counts included."""
    g = random_graph(3000, 24000, seed=31, hubs=2)
    params = dict(seed=5, init=1, ds_mode=mode, nodal_batch=40)
    eng = usap.Engine(g, slab_slack=1, **params)
    lab, st = eng.partition()
    o = orc.Oracle(g, threads=16, **params)
    o.set_labels(np.arange(g.N, dtype=np.int32), g.N)
    B, olab, ost = o.partition()
    assert st["error"] == 0 and ost["error"] == 0
    assert st["B"] == B and st["S"] == ost["S"]
    np.testing.assert_array_equal(lab, olab)
    assert (st["sweeps"], st["batches"]) == (ost["sweeps"], ost["batches"])
    assert [(p["B"], p["sweeps"]) for p in eng.partition_log()] == \
        [(p["B"], p["sweeps"]) for p in o.partition_log()]


def _oracle_sweep(o, N, batch, stream0, S, mean, thr, check):
    """The update step of graph_parallel.hpp:433-458 on the host over the
    oracle's batches: cumulative dS, S += dS_sum, the convergence test."""
    cum, inf, done, moves, trace = 0, False, 0, 0, []
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
        dS = -math.inf if inf else _fx_to_double(cum)
        S += dS
        trace.append(S)
        if check and -mean < thr * S:
            converged = True
            break
    return dict(batches=done, moves=moves, S=S, dS_sum=dS, converged=converged, trace=trace)


def test_sweep_stops_inside_a_chunk(usap, orc):
    """A sweep of 75 batches whose test fires at batch 41, inside the third
    launch chunk (batches 32-47), while the fourth is already queued: the
    device stops there, the batches launched after the stop leave the state
    alone, and labels, M and S equal the oracle's after exactly 41 batches.
    (The test inputs put the firing point mid-sweep: with thr < 0 the test
    -mean < thr * S fires once S falls below -mean / thr.  Exact dS mode, so
    no accepted dS is -inf.)"""
    g = random_graph(3000, 24000, seed=32, hubs=1)
    lab0 = np.arange(g.N, dtype=np.int32)
    probe = orc.Oracle(g, threads=16, seed=6, ds_mode=1)
    probe.set_labels(lab0, g.N)
    probe.block_merge(0)
    probe.carry_out(g.N // 2)
    _, S0 = probe.entropy()
    tr = _oracle_sweep(probe, g.N, 40, 0, S0, 0.0, 0.0, False)["trace"]
    j = 40  # fire after batch index 40 (the 41st)
    while tr[j] >= tr[j - 1]:
        j += 1
    level = 0.5 * (tr[j] + tr[j - 1])
    thr = -1e-6
    mean = level * 1e-6  # -mean < thr * S  <=>  S < level
    eng = usap.Engine(g, seed=6, init=1, ds_mode=1)
    eng.set_labels(lab0, g.N)
    eng.merge_propose(0)
    eng.merge_apply(g.N // 2)
    o = orc.Oracle(g, threads=16, seed=6, ds_mode=1)
    o.set_labels(lab0, g.N)
    o.block_merge(0)
    o.carry_out(g.N // 2)
    a = eng.nodal_sweep(40, 0, S0, mean, thr, 1)
    b = _oracle_sweep(o, g.N, 40, 0, S0, mean, thr, True)
    assert b["converged"] and b["batches"] == j + 1
    assert a["batches"] == b["batches"] and bool(a["converged"])
    assert a["moves"] == b["moves"] and a["S"] == b["S"]
    np.testing.assert_array_equal(eng.labels(), o.labels())
    np.testing.assert_array_equal(eng.M_dense(), o.M_dense())


def test_c3_full_size_state_invariants(usap):
    """The N=500k workload (identity start, 4096-node batches): after the
    first M build, after merge rounds down to B = 250000, 125000, 31250,
    3907 and 123 (per-node build, per-edge inserts, block-pair sums, LDS
    counts) and after two nodal sweeps at each (rows grown and compacted
    between batches), the device's M is the edge count of its labels, the
    degrees are M's row and column sums, and the MDL is the host recompute
    from that M, exactly."""
    g = usap.config_graph("c3")
    lam = usap.lambda_table(int(g.E))
    eng = usap.Engine(g, seed=3, init=1, nodal_batch=4096)
    eng.set_labels(np.arange(g.N, dtype=np.int32), g.N)
    _flush(f"c3 check B={eng.B}")
    _check_state(usap, eng, g, lam)
    check_at = {250000, 125000, 31250, 3907, 123}
    stream = 0
    for phase in range(13):
        eng.merge_propose(phase)
        eng.merge_apply(eng.B // 2)
        B = eng.B
        for _ in range(2):
            sw = eng.nodal_sweep(4096, stream, 0.0, 0.0, 0.0, 0)
            assert sw["batches"] == -(-g.N // 4096)
            stream += sw["batches"]
        if B in check_at:
            _flush(f"c3 check B={B}")
            _check_state(usap, eng, g, lam)


@pytest.mark.parametrize("B", [20000, 10000, 2000])
def test_row_build_large_B_weighted(usap, B):
    """The M builds on a weighted graph with self-loops, hubs and isolated
    nodes, with blocks of one node to thousands: per row (B >= 16384: small
    rows laid out in LDS, hashed rows above the LDS slab and a direct row
    inserted in place), edge by edge (B = 10000) and by sorted block pairs
    (B = 2000).  M, the degrees and the MDL equal the host recompute after
    the build and after a nodal sweep that moves nodes between such rows."""
    g = random_graph(40000, 400000, seed=21, weighted=True, self_loops=True, hubs=3,
                     isolated=50)
    lam = usap.lambda_table(int(g.out_w.sum()))
    lab = np.empty(g.N, np.int32)
    lab[:3000] = 0                                # a direct row (the hubs are here)
    lab[3000:3800] = 1 + np.arange(800) % 20      # 20 blocks of 40 nodes
    lab[3800:] = np.arange(g.N - 3800) % B        # every block used, most small
    eng = usap.Engine(g, seed=5, nodal_batch=4096)
    eng.set_labels(lab, B)
    _flush(f"row build B={eng.B}")
    _check_state(usap, eng, g, lam)
    sw = eng.nodal_sweep(4096, 0, 0.0, 0.0, 0.0, 0)
    assert sw["moves"] > 0
    _check_state(usap, eng, g, lam)
    eng.close()
