"""Python handle over one usap_engine (include/usap.h)."""
import ctypes as C

import numpy as np

from ._lib import (KERNELS, UsapKernelProf, UsapParams, UsapPhase, UsapStats, UsapSweep, check,
                   lib, ptr)

FRAC = 48
INIT_SCC, INIT_IDENTITY = 0, 1
DS_REF, DS_EXACT = 0, 1


def default_params(**kw):
    p = UsapParams()
    lib().usap_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def fx(hi, lo):
    """Exact int128 (scale 2^-48) from its (hi, lo) words."""
    return (int(hi) << 64) | int(np.uint64(lo))


class Engine:
    """One graph resident on the device: labels, M (block rows holding both orientations), block degrees."""

    def __init__(self, graph, edges=None, **params):
        """edges = (src, dst, w or None) in delivery order: the CSR/CSC are
        built on the device from that list (usap_create_from_edges) instead
        of uploaded from graph (which then only supplies N and the truth)."""
        self.graph = graph
        self.params = default_params(**params)
        h = C.c_void_p()
        if edges is not None:
            src, dst, w = edges
            check(lib().usap_create_from_edges(
                graph.N, len(src), ptr(np.ascontiguousarray(src, np.int32)),
                ptr(np.ascontiguousarray(dst, np.int32)),
                None if w is None else ptr(np.ascontiguousarray(w, np.int32)),
                C.byref(self.params), C.byref(h)))
        else:
            self._gs = graph.as_struct()
            check(lib().usap_create(C.byref(self._gs), C.byref(self.params), C.byref(h)))
        self.h = h.value

    def append_edges(self, src, dst, w=None):
        """Append edges after the current ones (usap_append_edges)."""
        src = np.ascontiguousarray(src, np.int32)
        dst = np.ascontiguousarray(dst, np.int32)
        wp = None if w is None else ptr(np.ascontiguousarray(w, np.int32))
        self._ck(lib().usap_append_edges(self.h, len(src), ptr(src), ptr(dst), wp))

    def close(self):
        if getattr(self, "h", None):
            lib().usap_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ck(self, rc):
        return check(rc, self.h)

    # ---- state ----------------------------------------------------------
    def set_labels(self, labels, B):
        self._ck(lib().usap_set_labels(self.h, np.ascontiguousarray(labels, np.int32), int(B)))

    def labels(self):
        out = np.zeros(self.graph.N, np.int32)
        self._ck(lib().usap_get_labels(self.h, out))
        return out

    @property
    def B(self):
        b = C.c_int32()
        self._ck(lib().usap_num_blocks(self.h, C.byref(b)))
        return b.value

    def degrees(self):
        B = self.B
        do, di = np.zeros(B, np.int32), np.zeros(B, np.int32)
        self._ck(lib().usap_get_degrees(self.h, do, di))
        return do, di

    def M_coo(self):
        nnz = C.c_int64()
        self._ck(lib().usap_get_M_coo(self.h, 0, None, None, None, C.byref(nnz)))
        n = nnz.value
        r, c, v = (np.zeros(max(n, 1), np.int32) for _ in range(3))
        self._ck(lib().usap_get_M_coo(self.h, n, ptr(r), ptr(c), ptr(v), C.byref(nnz)))
        return r[:n], c[:n], v[:n]

    def M_dense(self):
        B = self.B
        r, c, v = self.M_coo()
        M = np.zeros((B, B), np.int64)
        M[r, c] = v
        return M

    def entropy(self):
        S, hi, lo = C.c_double(), C.c_int64(), C.c_uint64()
        self._ck(lib().usap_entropy(self.h, C.byref(S), C.byref(hi), C.byref(lo)))
        return fx(hi.value, lo.value), S.value

    # ---- hot path -------------------------------------------------------
    def merge_propose(self, stream, K=None):
        K = self.params.K if K is None else K
        self._ck(lib().usap_merge_propose(self.h, K, stream, None, None))
        B = self.B
        s = np.zeros(B, np.int32)
        hi = np.zeros(B, np.int64)
        lo = np.zeros(B, np.uint64)
        self._ck(lib().usap_merge_result_fx(self.h, s, hi, lo))
        return s, [fx(h, l) for h, l in zip(hi, lo)]

    def merge_apply(self, B_to_merge):
        b = C.c_int32()
        self._ck(lib().usap_merge_apply(self.h, int(B_to_merge), C.byref(b)))
        return b.value

    def nodal_batch(self, beg, end, stream, detail=True):
        dS, mv = C.c_double(), C.c_int32()
        self._ck(lib().usap_nodal_batch(self.h, beg, end, stream, C.byref(dS), C.byref(mv)))
        hi, lo, inf = C.c_int64(), C.c_uint64(), C.c_int32()
        self._ck(lib().usap_nodal_sum_fx(self.h, C.byref(hi), C.byref(lo), C.byref(inf)))
        out = dict(sum=fx(hi.value, lo.value), sum_inf=inf.value, moves=mv.value,
                   dS_sum=dS.value)
        if detail:
            n = end - beg
            s = np.zeros(max(n, 1), np.int32)
            h = np.zeros(max(n, 1), np.int64)
            l = np.zeros(max(n, 1), np.uint64)
            fi = np.zeros(max(n, 1), np.int32)
            ac = np.zeros(max(n, 1), np.int32)
            self._ck(lib().usap_nodal_detail(self.h, s, h, l, fi, ac))
            out.update(s=s[:n], dS=[fx(a, b) for a, b in zip(h[:n], l[:n])], inf=fi[:n],
                       accept=ac[:n])
        return out

    def nodal_sweep(self, batch, stream0, S, mean, thr, check):
        """One nodal sweep with the update step on the device (usap_nodal_sweep)."""
        out = UsapSweep()
        self._ck(lib().usap_nodal_sweep(self.h, batch, stream0, S, mean, thr, int(check),
                                        C.byref(out)))
        d = out.as_dict()
        d["dS_fx"] = fx(d["dS_hi"], d["dS_lo"])
        return d

    def set_uniforms(self, u=None):
        """Draw from a pre-drawn uniform stream (usap_set_uniforms): u holds
        four uniforms in [0, 1) per draw; None returns to Philox."""
        if u is None:
            self._ck(lib().usap_set_uniforms(self.h, None, 0))
            return
        u = np.ascontiguousarray(u, np.float64).ravel()
        self._ck(lib().usap_set_uniforms(self.h, ptr(u), u.size))

    def uniforms_used(self):
        """Draws consumed from the pre-drawn stream (usap_uniforms_used)."""
        d = C.c_int64()
        self._ck(lib().usap_uniforms_used(self.h, C.byref(d)))
        return d.value

    def snapshot(self, slot, S=0.0):
        self._ck(lib().usap_snapshot(self.h, slot, S))

    def restore(self, slot):
        self._ck(lib().usap_restore(self.h, slot))

    def partition(self):
        """Full partition() through the host C++ driver.  Returns (labels, stats)."""
        lab = np.zeros(self.graph.N, np.int32)
        st = UsapStats()
        self._ck(lib().usap_partition(self.h, lab.ctypes.data, C.byref(st)))
        return lab, st.as_dict()

    def set_graph(self, graph):
        """Replace the edges (same N), keeping labels and B (usap_set_graph)."""
        gs = graph.as_struct()
        self._ck(lib().usap_set_graph(self.h, C.byref(gs)))
        self.graph, self._gs = graph, gs

    def resume(self, stream0):
        """Nodal phase at the current partition (usap_resume): (labels, stats)."""
        lab = np.zeros(self.graph.N, np.int32)
        st = UsapStats()
        self._ck(lib().usap_resume(self.h, stream0, lab.ctypes.data, C.byref(st)))
        return lab, st.as_dict()

    def partition_log(self):
        n = C.c_int32()
        self._ck(lib().usap_partition_log(self.h, 0, None, C.byref(n)))
        arr = (UsapPhase * max(n.value, 1))()
        self._ck(lib().usap_partition_log(self.h, n.value, arr, C.byref(n)))
        return [{k: getattr(a, k) for k, _ in UsapPhase._fields_} for a in arr[:n.value]]

    def profile(self, enable=-1):
        """enable=1 resets and starts kernel timing, 0 stops, -1 queries.
        Returns {kernel: {launches, ms, bytes, model_bytes}} accumulated so
        far (model_bytes: the reference algorithm's compulsory bytes for the
        same work, nodal and merge kernels)."""
        arr = (UsapKernelProf * len(KERNELS))()
        self._ck(lib().usap_profile(self.h, enable, arr))
        return {k: {"launches": a.launches, "ms": a.ms, "bytes": a.bytes,
                    "model_bytes": a.model_bytes}
                for k, a in zip(KERNELS, arr)}


def scc_init(graph, depth=20):
    """SCC initial partition (graph_parallel.hpp:514-603): (labels, B)."""
    gs = graph.as_struct()
    lab = np.zeros(graph.N, np.int32)
    B = C.c_int32()
    check(lib().usap_scc_init(C.byref(gs), depth, lab, C.byref(B)))
    return lab, B.value


def philox4x64(ctr, key):
    out = np.zeros(4, np.uint64)
    lib().usap_philox4x64(np.asarray(ctr, np.uint64), np.asarray(key, np.uint64), out)
    return out


def lambda_table(maxv):
    out = np.zeros(maxv + 1, np.int64)
    check(lib().usap_lambda_table(maxv, out))
    return out


def model_S(N, E, B):
    return lib().usap_model_S(N, E, B)


def golden_update(slots, cur, rate=0.5):
    """slots: [(B, S, id)] * 3 (large, med, small); cur: (B, S, id).
    Returns (rc, slots, next, B_to_merge)."""
    Bs = np.array([s[0] for s in slots], np.int64)
    Ss = np.array([s[1] for s in slots], np.float64)
    ids = np.array([s[2] for s in slots], np.int64)
    nB, nS, nid, btm = C.c_int64(), C.c_double(), C.c_int64(), C.c_int64()
    rc = lib().usap_golden_update(Bs, Ss, ids, cur[0], cur[1], cur[2], rate, C.byref(nB),
                                  C.byref(nS), C.byref(nid), C.byref(btm))
    return rc, list(zip(Bs.tolist(), Ss.tolist(), ids.tolist())), (nB.value, nS.value,
                                                                    nid.value), btm.value
