"""ctypes wrapper of the CPU oracle (oracle/usap_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "liborc.so")

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class OrcGraph(C.Structure):
    _fields_ = [
        ("N", C.c_int32),
        ("E", C.c_int64),
        ("out_ptr", C.c_void_p),
        ("out_idx", C.c_void_p),
        ("in_ptr", C.c_void_p),
        ("in_idx", C.c_void_p),
        ("out_w", C.c_void_p),
        ("in_w", C.c_void_p),
    ]


class OrcParams(C.Structure):
    _fields_ = [
        ("K", C.c_int32),
        ("reduction_rate", C.c_double),
        ("max_nodal_itr", C.c_int32),
        ("num_batch_nodal_update", C.c_int32),
        ("nodal_batch", C.c_int32),
        ("thr1", C.c_double),
        ("thr2", C.c_double),
        ("avg_window", C.c_int32),
        ("dfs_depth", C.c_int32),
        ("block_size", C.c_int32),
        ("init", C.c_int32),
        ("ds_mode", C.c_int32),
        ("seed", C.c_uint64),
        ("threads", C.c_int32),
    ]


class OrcPhase(C.Structure):
    _fields_ = [("B", C.c_int64), ("S", C.c_double), ("sweeps", C.c_int64),
                ("batches", C.c_int64), ("moves", C.c_int64)]


class OrcSnap(C.Structure):
    _fields_ = [("B", C.c_int64), ("S", C.c_double), ("id", C.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.orc_default_params.argtypes = [C.POINTER(OrcParams)]
        L.orc_philox4x64.argtypes = [_u64p, _u64p, _u64p]
        L.orc_lambda_table.argtypes = [C.c_int64, _i64p]
        L.orc_pick.argtypes = [_i64p, C.c_int64, C.c_int64]
        L.orc_pick.restype = C.c_int64
        L.orc_scc_init.argtypes = [C.POINTER(OrcGraph), C.c_int32, _i32p]
        L.orc_scc_init.restype = C.c_int32
        L.orc_create.argtypes = [C.POINTER(OrcGraph), C.POINTER(OrcParams)]
        L.orc_create.restype = C.c_void_p
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_labels.argtypes = [C.c_void_p, _i32p, C.c_int32]
        L.orc_num_blocks.argtypes = [C.c_void_p]
        L.orc_num_blocks.restype = C.c_int32
        L.orc_get_labels.argtypes = [C.c_void_p, _i32p]
        L.orc_get_degrees.argtypes = [C.c_void_p, _i64p, _i64p]
        L.orc_get_M_dense.argtypes = [C.c_void_p, _i64p]
        L.orc_entropy.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_double)]
        L.orc_merge_propose_one.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_uint64,
                                            C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                            C.POINTER(C.c_uint64)]
        L.orc_block_merge.argtypes = [C.c_void_p, C.c_uint64, _i32p, _i64p, _u64p]
        L.orc_block_merge.restype = C.c_int
        L.orc_carry_out.argtypes = [C.c_void_p, C.c_int32]
        L.orc_nodal_propose_one.argtypes = [C.c_void_p, C.c_int32, C.c_uint64,
                                            C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                            C.POINTER(C.c_uint64), C.POINTER(C.c_int32),
                                            C.POINTER(C.c_int32)]
        L.orc_nodal_batch.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_uint64,
                                      _i32p, _i64p, _u64p, _i32p, _i32p,
                                      C.POINTER(C.c_int64), C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.orc_nodal_batch.restype = C.c_int
        L.orc_set_uniforms.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        L.orc_set_uniforms.restype = C.c_int
        L.orc_uniforms_used.argtypes = [C.c_void_p]
        L.orc_uniforms_used.restype = C.c_int64
        L.orc_nodal_eval.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_uint64,
                                     C.POINTER(C.c_int32)]
        L.orc_merge_eval.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_uint64]
        L.orc_merge_eval.restype = None
        L.orc_golden_update.argtypes = [C.POINTER(OrcSnap), OrcSnap, C.c_double,
                                        C.POINTER(OrcSnap), C.POINTER(C.c_int64)]
        L.orc_partition.argtypes = [C.c_void_p, _i32p, _f64p]
        L.orc_partition.restype = C.c_int32
        L.orc_partition_log.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.orc_partition_log.restype = C.c_int32
        L.orc_nodal_phase_scripted.argtypes = [C.POINTER(OrcParams), C.c_int32, C.c_int32,
                                               C.c_int32, _f64p, C.c_int64, _i64p, _u64p,
                                               _i32p, _i32p, C.c_int64, _f64p, _f64p]
        L.orc_nodal_phase_scripted.restype = C.c_int
        L.orc_data_S_double.argtypes = [C.c_void_p]
        L.orc_data_S_double.restype = C.c_double
        L.orc_model_S.argtypes = [C.c_int64, C.c_int64, C.c_int64]
        L.orc_model_S.restype = C.c_double
        _lib = L
    return _lib


FRAC = 48


def fx(hi, lo):
    """int128 fixed point (scale 2^-48) from its two words, as a Python int."""
    return (int(hi) << 64) | int(lo)


def fx_to_float(v):
    return float(v) / float(1 << FRAC)


def philox4x64(ctr, key):
    out = np.zeros(4, np.uint64)
    lib().orc_philox4x64(np.asarray(ctr, np.uint64), np.asarray(key, np.uint64), out)
    return out


def lambda_table(maxv):
    out = np.zeros(maxv + 1, np.int64)
    lib().orc_lambda_table(maxv, out)
    return out


def pick(w, t):
    return lib().orc_pick(np.ascontiguousarray(w, np.int64), len(w), int(t))


def default_params(**kw):
    p = OrcParams()
    lib().orc_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def golden_update(slots, cur, rate=0.5):
    """slots: list of 3 (B, S, id) tuples (large, med, small); cur: (B, S, id).
    Returns (rc, new slots, next (B, S, id), B_to_merge)."""
    arr = (OrcSnap * 3)(*[OrcSnap(*s) for s in slots])
    nxt = OrcSnap(0, 0.0, -1)
    btm = C.c_int64(0)
    rc = lib().orc_golden_update(arr, OrcSnap(*cur), rate, C.byref(nxt), C.byref(btm))
    return rc, [(a.B, a.S, a.id) for a in arr], (nxt.B, nxt.S, nxt.id), btm.value


def nodal_phase_scripted(N, batch, isf, S_start, sums, moves=None, **params):
    """The oracle's nodal-phase loop over scripted inputs: S_start[i] = MDL at
    sweep i's start; sums[j] = batch j's accepted-dS sum (a Python int in
    fixed point, or -math.inf).  Returns (stats dict, itr_dS)."""
    p = default_params(**params)
    hi = np.array([0 if s == -math.inf else (int(s) >> 64) for s in sums], np.int64)
    lo = np.array([0 if s == -math.inf else (int(s) & ((1 << 64) - 1)) for s in sums], np.uint64)
    inf = np.array([-1 if s == -math.inf else 0 for s in sums], np.int32)
    mv = np.ascontiguousarray(moves if moves is not None else np.zeros(len(sums)), np.int32)
    out = np.zeros(6, np.float64)
    itr = np.zeros(max(p.max_nodal_itr, 1), np.float64)
    rc = lib().orc_nodal_phase_scripted(C.byref(p), N, batch, isf,
                                        np.ascontiguousarray(S_start, np.float64), len(S_start),
                                        hi, lo, inf, mv, len(sums), out, itr)
    if rc != 0:
        raise ValueError("script ran out")
    keys = ["sweeps", "batches", "proposals", "moves", "converged", "S_last"]
    return dict(zip(keys, out.tolist())), itr


class Oracle:
    """One oracle state over a CSR/CSC graph (see usap_amd.graph.Graph)."""

    def __init__(self, g, **params):
        self._keep = []
        self.g = g
        og = OrcGraph()
        og.N = g.N
        og.E = g.E
        for name in ("out_ptr", "out_idx", "in_ptr", "in_idx"):
            a = np.ascontiguousarray(getattr(g, name))
            self._keep.append(a)
            setattr(og, name, a.ctypes.data)
        for name in ("out_w", "in_w"):
            a = getattr(g, name)
            if a is None:
                setattr(og, name, None)
            else:
                a = np.ascontiguousarray(a, np.int32)
                self._keep.append(a)
                setattr(og, name, a.ctypes.data)
        self._og = og
        self.params = default_params(**params)
        self.st = lib().orc_create(C.byref(og), C.byref(self.params))

    def close(self):
        if self.st:
            lib().orc_destroy(self.st)
            self.st = None

    def __del__(self):
        self.close()

    def scc(self, depth=20):
        lab = np.zeros(self.g.N, np.int32)
        B = lib().orc_scc_init(C.byref(self._og), depth, lab)
        return lab, B

    def set_labels(self, labels, B):
        rc = lib().orc_set_labels(self.st, np.ascontiguousarray(labels, np.int32), int(B))
        if rc != 0:
            raise ValueError("bad labels")

    @property
    def B(self):
        return lib().orc_num_blocks(self.st)

    def labels(self):
        out = np.zeros(self.g.N, np.int32)
        lib().orc_get_labels(self.st, out)
        return out

    def degrees(self):
        B = self.B
        do, di = np.zeros(B, np.int64), np.zeros(B, np.int64)
        lib().orc_get_degrees(self.st, do, di)
        return do, di

    def M_dense(self):
        B = self.B
        out = np.zeros(B * B, np.int64)
        lib().orc_get_M_dense(self.st, out)
        return out.reshape(B, B)

    def entropy(self):
        hi, lo, S = C.c_int64(), C.c_uint64(), C.c_double()
        lib().orc_entropy(self.st, C.byref(hi), C.byref(lo), C.byref(S))
        return fx(hi.value, lo.value), S.value

    def data_S_double(self):
        return lib().orc_data_S_double(self.st)

    def merge_propose_one(self, r, idx, stream):
        s, hi, lo = C.c_int32(), C.c_int64(), C.c_uint64()
        lib().orc_merge_propose_one(self.st, r, idx, stream, C.byref(s), C.byref(hi), C.byref(lo))
        return s.value, fx(hi.value, lo.value)

    def block_merge(self, stream):
        B = self.B
        s = np.zeros(B, np.int32)
        hi = np.zeros(B, np.int64)
        lo = np.zeros(B, np.uint64)
        if lib().orc_block_merge(self.st, stream, s, hi, lo):
            raise RuntimeError("orc_block_merge: pre-drawn uniform stream exhausted")
        return s, [fx(h, l) for h, l in zip(hi, lo)]

    def carry_out(self, B_to_merge):
        return lib().orc_carry_out(self.st, int(B_to_merge))

    def nodal_propose_one(self, n, stream):
        s, hi, lo, inf, acc = C.c_int32(), C.c_int64(), C.c_uint64(), C.c_int32(), C.c_int32()
        lib().orc_nodal_propose_one(self.st, n, stream, C.byref(s), C.byref(hi), C.byref(lo),
                                    C.byref(inf), C.byref(acc))
        return s.value, fx(hi.value, lo.value), inf.value, acc.value

    def nodal_batch(self, beg, end, stream):
        n = end - beg
        s = np.zeros(max(n, 1), np.int32)
        hi = np.zeros(max(n, 1), np.int64)
        lo = np.zeros(max(n, 1), np.uint64)
        inf = np.zeros(max(n, 1), np.int32)
        acc = np.zeros(max(n, 1), np.int32)
        shi, slo, sinf, mv = C.c_int64(), C.c_uint64(), C.c_int32(), C.c_int32()
        if lib().orc_nodal_batch(self.st, beg, end, stream, s, hi, lo, inf, acc, C.byref(shi),
                                 C.byref(slo), C.byref(sinf), C.byref(mv)):
            raise RuntimeError("orc_nodal_batch: pre-drawn uniform stream exhausted")
        dS = [fx(h, l) for h, l in zip(hi[:n], lo[:n])]
        return dict(s=s[:n].copy(), dS=dS, inf=inf[:n].copy(), accept=acc[:n].copy(),
                    sum=fx(shi.value, slo.value), sum_inf=sinf.value, moves=mv.value)

    def set_uniforms(self, u=None):
        """Draw from a pre-drawn uniform stream (four uniforms in [0, 1) per
        draw); None returns to Philox."""
        if u is None:
            rc = lib().orc_set_uniforms(self.st, None, 0)
        else:
            u = np.ascontiguousarray(u, np.float64).ravel()
            self._uni = u
            rc = lib().orc_set_uniforms(self.st, u.ctypes.data, u.size)
        if rc:
            raise ValueError("orc_set_uniforms: n must be a multiple of 4, values in [0, 1)")

    def uniforms_used(self):
        return lib().orc_uniforms_used(self.st)

    def nodal_eval(self, beg, end, stream):
        a = C.c_int32()
        lib().orc_nodal_eval(self.st, beg, end, stream, C.byref(a))
        return a.value

    def merge_eval(self, r0, r1, stream):
        """The K merge proposals of blocks [r0, r1) (timing only)."""
        lib().orc_merge_eval(self.st, r0, r1, stream)

    def partition(self):
        lab = np.zeros(self.g.N, np.int32)
        stats = np.zeros(8, np.float64)
        B = lib().orc_partition(self.st, lab, stats)
        keys = ["S", "phases", "sweeps", "batches", "nodal_proposals", "moves",
                "merge_proposals", "error"]
        return B, lab, dict(zip(keys, stats.tolist()))

    def partition_log(self):
        """Per-phase records of the last partition(): B, S, sweeps, batches, moves."""
        n = lib().orc_partition_log(self.st, 0, None)
        arr = (OrcPhase * max(n, 1))()
        lib().orc_partition_log(self.st, n, arr)
        return [{k: getattr(a, k) for k, _ in OrcPhase._fields_} for a in arr[:n]]
