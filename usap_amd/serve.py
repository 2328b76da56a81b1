"""Serving graphs that arrive as streamed edge increments (usap_serve in
include/usap.h) and the served workload: graphs with N log-uniform in
[1000, 200000] (the batched workload's generator settings), Poisson
arrivals, each graph's edges shuffled and delivered in 10 increments."""
import ctypes as C

import numpy as np

from ._lib import (UsapError, UsapParams, UsapServeGraph, UsapServeOpts, UsapServeStage,
                   UsapServeStats, lib)
from .engine import default_params
from .graph import generate_dcsbm


class ServedGraph:
    """One request: N, the edges in delivery order, the planted truth."""

    def __init__(self, N, src, dst, truth=None, w=None, seed=0):
        self.N = int(N)
        self.src = np.ascontiguousarray(src, np.int32)
        self.dst = np.ascontiguousarray(dst, np.int32)
        self.w = None if w is None else np.ascontiguousarray(w, np.int32)
        self.truth = truth
        self.seed = seed

    @property
    def E(self):
        return len(self.src)


def served_graphs(n=32, seed=5, n_min=1000, n_max=200000):
    """n requests of the served workload; graph i's generator and shuffle use
    seed*100000 + i."""
    rng = np.random.default_rng(seed)
    Ns = np.rint(np.exp(rng.uniform(np.log(n_min), np.log(n_max), n))).astype(int)
    p_ins = rng.uniform(0.55, 0.85, n)
    out = []
    for i, (N, p_in) in enumerate(zip(Ns, p_ins)):
        gseed = seed * 100000 + i
        g = generate_dcsbm(int(N), max(2, int(round(N ** 0.4))), 3.0, 20, 10, 100, 2.5,
                           float(p_in), gseed)
        src, dst = g.edges()
        perm = np.random.default_rng(gseed).permutation(len(src))
        out.append(ServedGraph(g.N, src[perm], dst[perm], g.truth, seed=gseed))
    return out


def poisson_arrivals(n, rate, seed=5):
    """Arrival times (s) of a Poisson process of `rate` graphs/s, first at 0."""
    gaps = np.random.default_rng(seed + 17).exponential(1.0 / rate, n)
    gaps[0] = 0.0
    return np.cumsum(gaps)


def serve(graphs, arrivals, params, stages=10, interval_s=0.0, workers=4, hbm_budget=0,
          final_partition=False, repartition_growth=0.0):
    """Serves the graphs (params: dict for all or one per graph).  Returns
    (final labels per graph, stage records [graph][stage], stats dict).
    repartition_growth > 1: a stage whose edges reach that multiple of the
    edges at the graph's last full partition partitions afresh."""
    n = len(graphs)
    arr = (UsapServeGraph * max(n, 1))()
    for i, g in enumerate(graphs):
        p = params[i] if isinstance(params, (list, tuple)) else params
        arr[i].N = g.N
        arr[i].E = g.E
        arr[i].src = g.src.ctypes.data
        arr[i].dst = g.dst.ctypes.data
        arr[i].w = None if g.w is None else g.w.ctypes.data
        arr[i].arrival_s = float(arrivals[i])
        arr[i].params = p if isinstance(p, UsapParams) else default_params(**p)
    opt = UsapServeOpts(stages, float(interval_s), workers, int(hbm_budget), int(final_partition),
                        float(repartition_growth))
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([g.N for g in graphs])
    labels = np.zeros(max(int(off[-1]), 1), np.int32)
    recs = (UsapServeStage * max(n * stages, 1))()
    st = UsapServeStats()
    rc = lib().usap_serve(n, arr, C.byref(opt), labels.ctypes.data, off.ctypes.data, recs,
                          C.byref(st))
    stats = st.as_dict()
    stats["rc"] = rc
    stats["error"] = lib().usap_last_error(None).decode() if rc != 0 else ""
    if rc != 0 and stats["failed"] == 0:
        raise UsapError(rc, stats["error"])
    per = [[recs[i * stages + k].as_dict() for k in range(stages)] for i in range(n)]
    return [labels[off[i]:off[i + 1]].copy() for i in range(n)], per, stats


def footprint(N, E):
    return lib().usap_footprint(int(N), int(E))
