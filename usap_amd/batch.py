"""Many graphs at once: usap_partition_batch / usap_batch_autotune
(include/usap.h) and the batched DCSBM workload (256 graphs, N log-uniform
in [500, 50000], round(N^0.4) planted blocks, within-block fraction uniform
in [0.55, 0.85], otherwise the N=1000 generator settings)."""
import ctypes as C

import numpy as np

from ._lib import UsapBatchStats, UsapError, UsapGraph, UsapParams, UsapStats, lib
from .engine import default_params
from .graph import generate_dcsbm


def batch_graphs(n=256, seed=4, n_min=500, n_max=50000):
    """The batched workload: n seeded DCSBM graphs (graph i uses generator
    seed seed*100000 + i).  Returns (graphs, specs)."""
    rng = np.random.default_rng(seed)
    Ns = np.rint(np.exp(rng.uniform(np.log(n_min), np.log(n_max), n))).astype(int)
    p_ins = rng.uniform(0.55, 0.85, n)
    graphs, specs = [], []
    for i, (N, p_in) in enumerate(zip(Ns, p_ins)):
        K = max(2, int(round(N ** 0.4)))
        spec = dict(N=int(N), K=K, alpha=3.0, min_block=20, dmin=10, dmax=100, gamma=2.5,
                    p_in=float(p_in), seed=seed * 100000 + i)
        graphs.append(generate_dcsbm(**spec))
        specs.append(spec)
    return graphs, specs


def _arrays(graphs, params):
    n = len(graphs)
    gs = (UsapGraph * max(n, 1))()
    for i, g in enumerate(graphs):
        gs[i] = g.as_struct()
    ps = (UsapParams * max(n, 1))()
    for i in range(n):
        p = params[i] if isinstance(params, (list, tuple)) else params
        ps[i] = p if isinstance(p, UsapParams) else default_params(**p)
    return gs, ps


def partition_batch(graphs, params, workers=8):
    """Partitions every graph (params: one dict for all, or one per graph),
    `workers` graphs in flight (8, the autotuner's usual pick on one MI355X,
    by default; 0 = autotuned).  Returns (labels list, stats list, batch
    stats dict)."""
    n = len(graphs)
    gs, ps = _arrays(graphs, params)
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([g.N for g in graphs])
    labels = np.zeros(max(int(off[-1]), 1), np.int32)
    stats = (UsapStats * max(n, 1))()
    bs = UsapBatchStats()
    rc = lib().usap_partition_batch(n, gs, ps, labels.ctypes.data, off.ctypes.data, stats,
                                    workers, C.byref(bs))
    out = bs.as_dict()
    st = [stats[i].as_dict() for i in range(n)]
    out["rc"] = rc
    out["error"] = lib().usap_last_error(None).decode() if rc != 0 else ""
    if rc != 0 and out["failed"] == 0:
        raise UsapError(rc, out["error"])
    return [labels[off[i]:off[i + 1]].copy() for i in range(n)], st, out


def autotune(graphs, params, max_workers=0, sample=0):
    """Workers chosen by usap_batch_autotune, with the trials (workers, graphs/s)."""
    n = len(graphs)
    gs, ps = _arrays(graphs, params)
    w = C.c_int32()
    tried = np.zeros(16, np.int32)
    rates = np.zeros(16, np.float64)
    rc = lib().usap_batch_autotune(n, gs, ps, max_workers, sample, C.byref(w),
                                   tried.ctypes.data, rates.ctypes.data)
    if rc != 0:
        raise UsapError(rc, lib().usap_last_error(None).decode())
    trials = [(int(a), float(b)) for a, b in zip(tried, rates) if a > 0]
    return w.value, trials
