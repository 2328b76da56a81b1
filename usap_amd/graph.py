"""Graphs: CSR/CSC in the reference's neighbour order, the GraphChallenge TSV
loader and the seeded DCSBM generator (all through libusap.so host code)."""
import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from ._lib import UsapGraph, check, lib, ptr


@dataclass
class Graph:
    N: int
    E: int
    out_ptr: np.ndarray  # int64 [N+1]
    out_idx: np.ndarray  # int32 [E]
    in_ptr: np.ndarray   # int64 [N+1]
    in_idx: np.ndarray   # int32 [E]
    out_w: Optional[np.ndarray] = None  # int32 [E] or None (unweighted)
    in_w: Optional[np.ndarray] = None
    truth: Optional[np.ndarray] = None  # planted / true partition, 0-based

    @staticmethod
    def from_edges(src, dst, N=None, w=None, truth=None):
        """CSR (out) / CSC (in) from an edge list in file order; both lists
        keep file order, as graph_parallel.hpp:290-294 builds them."""
        src = np.ascontiguousarray(src, np.int32)
        dst = np.ascontiguousarray(dst, np.int32)
        E = len(src)
        if N is None:
            N = int(max(src.max(initial=-1), dst.max(initial=-1)) + 1)
        if w is not None:
            w = np.ascontiguousarray(w, np.int32)
            if np.all(w == 1):
                w = None
        out_ptr = np.zeros(N + 1, np.int64)
        in_ptr = np.zeros(N + 1, np.int64)
        out_idx = np.zeros(max(E, 1), np.int32)
        in_idx = np.zeros(max(E, 1), np.int32)
        out_w = np.zeros(max(E, 1), np.int32) if w is not None else None
        in_w = np.zeros(max(E, 1), np.int32) if w is not None else None
        check(lib().usap_build_csr(N, E, ptr(src), ptr(dst), ptr(w), ptr(out_ptr), ptr(out_idx),
                                   ptr(out_w), ptr(in_ptr), ptr(in_idx), ptr(in_w)))
        return Graph(N, E, out_ptr, out_idx[:E], in_ptr, in_idx[:E],
                     None if w is None else out_w[:E], None if w is None else in_w[:E],
                     None if truth is None else np.ascontiguousarray(truth, np.int32))

    def as_struct(self):
        """usap_graph view (keeps references alive through self)."""
        g = UsapGraph()
        g.N = self.N
        g.E = self.E
        self._keep = [np.ascontiguousarray(a) for a in
                      (self.out_ptr, self.out_idx, self.in_ptr, self.in_idx)]
        g.out_ptr, g.out_idx, g.in_ptr, g.in_idx = (a.ctypes.data for a in self._keep)
        if self.out_w is not None:
            self._keep += [np.ascontiguousarray(self.out_w, np.int32),
                           np.ascontiguousarray(self.in_w, np.int32)]
            g.out_w = self._keep[-2].ctypes.data
            g.in_w = self._keep[-1].ctypes.data
        return g

    def edges(self):
        """(src, dst) in CSR order."""
        src = np.repeat(np.arange(self.N, dtype=np.int32), np.diff(self.out_ptr))
        return src, self.out_idx.copy()


def load_tsv(path, true_path=None):
    """GraphChallenge TSV ("from\\tto\\tweight", 1-based) -> Graph.
    load_graph_from_tsv (graph_parallel.hpp:251-317), except that N is the
    largest id in either column (the reference uses the largest source id)."""
    N, E = C.c_int32(), C.c_int64()
    check(lib().usap_load_tsv(path.encode(), C.byref(N), C.byref(E), None, None, None))
    src = np.zeros(max(E.value, 1), np.int32)
    dst = np.zeros(max(E.value, 1), np.int32)
    w = np.zeros(max(E.value, 1), np.int32)
    check(lib().usap_load_tsv(path.encode(), C.byref(N), C.byref(E), ptr(src), ptr(dst), ptr(w)))
    truth = None
    if true_path is not None:
        t = np.loadtxt(true_path, dtype=np.int64, ndmin=2)
        truth = np.zeros(N.value, np.int32)
        truth[t[:, 0] - 1] = t[:, 1] - 1
    E = E.value
    return Graph.from_edges(src[:E], dst[:E], N.value, w[:E], truth)


def generate_dcsbm(N, K, alpha, min_block, dmin, dmax, gamma=2.5, p_in=0.8, seed=0):
    """Seeded directed degree-corrected SBM (see usap_generate_dcsbm)."""
    E = C.c_int64()
    args = (N, K, float(alpha), min_block, dmin, dmax, float(gamma), float(p_in), seed)
    check(lib().usap_generate_dcsbm(*args, C.byref(E), None, None, None))
    src = np.zeros(max(E.value, 1), np.int32)
    dst = np.zeros(max(E.value, 1), np.int32)
    truth = np.zeros(N, np.int32)
    check(lib().usap_generate_dcsbm(*args, C.byref(E), ptr(src), ptr(dst), ptr(truth)))
    E = E.value
    return Graph.from_edges(src[:E], dst[:E], N, None, truth)


# The synthetic workloads: N, planted blocks, Dirichlet alpha, min block size,
# out-degree range, power-law exponent, within-block edge fraction, seed.
CONFIGS = {
    "c1": dict(N=1000, K=11, alpha=3.0, min_block=20, dmin=10, dmax=100, gamma=2.5, p_in=0.8,
               seed=1),
    "c2": dict(N=20000, K=32, alpha=0.8, min_block=50, dmin=10, dmax=100, gamma=2.5, p_in=0.6,
               seed=2),
    "c3": dict(N=500000, K=120, alpha=3.0, min_block=20, dmin=10, dmax=200, gamma=2.5,
               p_in=0.7, seed=3),
}


def config_graph(name, **over):
    kw = dict(CONFIGS[name])
    kw.update(over)
    return generate_dcsbm(**kw)
