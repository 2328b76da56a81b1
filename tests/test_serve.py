"""Serving loop (usap_serve): graphs arriving over time as streamed edge
increments.  CPU: argument checks, footprint, the workload generator.  GPU:
each graph's per-stage results equal those it gets served alone, stage 1 is
a full partition of the first increment and later stages resume the chain
on the grown graph, and the footprint budget bounds the resident graphs."""
import numpy as np
import pytest

from usap_amd import serve as sv


def test_footprint_and_workload(usap):
    assert sv.footprint(1000, 20000) < sv.footprint(2000, 20000) < sv.footprint(2000, 40000)
    assert sv.footprint(0, 10) == 0
    gs = sv.served_graphs(n=4, seed=5, n_min=1000, n_max=3000)
    again = sv.served_graphs(n=4, seed=5, n_min=1000, n_max=3000)
    for g, h in zip(gs, again):
        assert 1000 <= g.N <= 3000 and g.E > 10 * g.N
        np.testing.assert_array_equal(g.src, h.src)
        assert not np.all(np.diff(g.src) >= 0)  # delivery order is shuffled
    t = sv.poisson_arrivals(100, rate=10.0)
    assert t[0] == 0 and np.all(np.diff(t) >= 0) and 5 < t[-1] < 20


def test_serve_argument_checks(usap):
    import ctypes as C
    from usap_amd import _lib
    L = _lib.lib()
    st = _lib.UsapServeStats()
    opt = _lib.UsapServeOpts(10, 0.0, 2, 0)
    assert L.usap_serve(0, None, C.byref(opt), None, None, None, C.byref(st)) == 0
    for bad in (_lib.UsapServeOpts(0, 0.0, 2, 0), _lib.UsapServeOpts(10, -1.0, 2, 0),
                _lib.UsapServeOpts(10, 0.0, 0, 0), _lib.UsapServeOpts(10, 0.0, 2, -5)):
        assert L.usap_serve(0, None, C.byref(bad), None, None, None, None) == _lib.USAP_EINVAL
    assert L.usap_serve(1, None, C.byref(opt), None, None, None, None) == _lib.USAP_EINVAL
    # captured-sweep replay with several stages in flight is refused before
    # any device call
    gs = sv.served_graphs(n=2, seed=8, n_min=300, n_max=400)
    with pytest.raises(_lib.UsapError) as ei:
        sv.serve(gs, [0.0, 0.0], dict(seed=1, graphs=1), stages=2, workers=2)
    assert ei.value.code == _lib.USAP_EINVAL and "workers=1" in str(ei.value)


def _small():
    gs = sv.served_graphs(n=5, seed=8, n_min=300, n_max=2500)
    params = [dict(seed=g.seed, init=i % 2, ds_mode=(i // 2) % 2) for i, g in enumerate(gs)]
    return gs, params


@pytest.mark.gpu
def test_serve_equals_alone(usap):
    gs, params = _small()
    t = sv.poisson_arrivals(len(gs), rate=40.0, seed=8)
    labs, per, st = sv.serve(gs, t, params, stages=4, interval_s=0.005, workers=3)
    assert st["failed"] == 0 and st["p99_ms"] >= st["p50_ms"] > 0
    for g, p, lab, recs in zip(gs, params, labs, per):
        assert [r["stage"] for r in recs] == [1, 2, 3, 4]
        assert all(r["t_done"] >= r["t_start"] >= r["t_ready"] - 1e-3 for r in recs)
        assert [r["E"] for r in recs] == [int(np.ceil(g.E * k / 4)) for k in (1, 2, 3, 4)]
        lab1, per1, st1 = sv.serve([g], [0.0], [p], stages=4, interval_s=0.0, workers=1)
        np.testing.assert_array_equal(lab, lab1[0])
        assert [r["S"] for r in recs] == [r["S"] for r in per1[0]]
        assert [r["B"] for r in recs] == [r["B"] for r in per1[0]]


@pytest.mark.gpu
def test_serve_stages_are_partition_then_resume(usap):
    gs, params = _small()
    g, p = gs[1], params[1]
    labs, per, st = sv.serve([g], [0.0], [p], stages=3, workers=1)
    eng = None
    for k in (1, 2, 3):
        Ek = int(np.ceil(g.E * k / 3))
        gk = usap.Graph.from_edges(g.src[:Ek], g.dst[:Ek], N=g.N)
        if k == 1:
            eng = usap.Engine(gk, **p)
            lab, stt = eng.partition()
        else:
            eng.set_graph(gk)
            lab, stt = eng.resume(k << 40)
        assert stt["S"] == per[0][k - 1]["S"] and stt["B"] == per[0][k - 1]["B"]
    np.testing.assert_array_equal(lab, labs[0])
    eng.close()


@pytest.mark.gpu
def test_serve_budget_bounds_resident_graphs(usap):
    gs, params = _small()
    # no two graphs fit together; a graph larger than the budget is admitted
    # only when nothing else is resident
    budget = int(1.5 * min(sv.footprint(g.N, g.E) for g in gs))
    t = np.zeros(len(gs))  # all arrive at once
    labs, per, st = sv.serve(gs, t, params, stages=2, workers=4, hbm_budget=budget)
    assert st["failed"] == 0 and st["max_resident"] == 1
    spans = sorted((r[0]["t_admit"], r[0]["t_release"]) for r in per)
    assert all(a[1] <= b[0] + 1e-6 for a, b in zip(spans, spans[1:]))
    labs2, per2, _ = sv.serve(gs, t, params, stages=2, workers=4)
    for a, b in zip(labs, labs2):
        np.testing.assert_array_equal(a, b)


@pytest.mark.gpu
def test_serve_final_partition(usap):
    """With final_partition the last stage is a full partition of the whole
    graph on the resident engine: the same as partitioning it alone."""
    gs, params = _small()
    g, p = gs[2], params[2]
    labs, per, st = sv.serve([g], [0.0], [p], stages=3, workers=1, final_partition=True)
    assert st["failed"] == 0
    full = usap.Graph.from_edges(g.src, g.dst, N=g.N)
    eng = usap.Engine(full, **p)
    lab, stt = eng.partition()
    eng.close()
    np.testing.assert_array_equal(lab, labs[0])
    assert stt["S"] == per[0][2]["S"] and stt["B"] == per[0][2]["B"]


@pytest.mark.gpu
def test_serve_repartition_on_growth(usap):
    """repartition_growth 2: a stage whose edges reach twice those of the
    last full partition partitions afresh (stages 1, 2, 4, 8 of 10 here);
    each such stage equals a full partition of its prefix alone, and the
    stages between resume the chain."""
    gs, params = _small()
    g, p = gs[3], params[3]
    labs, per, st = sv.serve([g], [0.0], [p], stages=10, workers=1, repartition_growth=2.0)
    assert st["failed"] == 0
    E = [int(np.ceil(g.E * k / 10)) for k in range(1, 11)]
    want, last = [], None
    for k, Ek in enumerate(E, 1):
        full = k == 1 or Ek >= 2 * last
        want.append(int(full))
        if full:
            last = Ek
    assert [r["full"] for r in per[0]] == want
    for k, rec in enumerate(per[0], 1):
        if rec["full"]:
            gk = usap.Graph.from_edges(g.src[:E[k - 1]], g.dst[:E[k - 1]], N=g.N)
            with usap.Engine(gk, **p) as eng:
                lab, stt = eng.partition()
            assert stt["S"] == rec["S"] and stt["B"] == rec["B"], k
