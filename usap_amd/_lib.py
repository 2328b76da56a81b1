"""ctypes bindings of the C-ABI in include/usap.h (usap_amd/lib/libusap.so).

The library is the product: if it is missing, every entry point raises.  There
is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libusap.so")

USAP_OK = 0
USAP_EINVAL, USAP_ENOMEM, USAP_EHIP, USAP_ESTATE, USAP_EINFEASIBLE, USAP_EIO = range(-1, -7, -1)
ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EHIP", -4: "ESTATE", -5: "EINFEASIBLE", -6: "EIO"}

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = C.c_void_p


class UsapError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__(f"usap error {code} ({ERRORS.get(code, '?')}): {msg}")


class UsapGraph(C.Structure):
    _fields_ = [
        ("N", C.c_int32),
        ("E", C.c_int64),
        ("out_ptr", _vp),
        ("out_idx", _vp),
        ("in_ptr", _vp),
        ("in_idx", _vp),
        ("out_w", _vp),
        ("in_w", _vp),
    ]


class UsapParams(C.Structure):
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
        ("init", C.c_int32),
        ("ds_mode", C.c_int32),
        ("seed", C.c_uint64),
        ("device", C.c_int32),
        ("verbose", C.c_int32),
        ("slab_slack", C.c_int32),
        ("graphs", C.c_int32),
    ]


class UsapStats(C.Structure):
    _fields_ = [
        ("S", C.c_double),
        ("B", C.c_int32),
        ("phases", C.c_int32),
        ("sweeps", C.c_int64),
        ("batches", C.c_int64),
        ("nodal_proposals", C.c_int64),
        ("accepted_moves", C.c_int64),
        ("merge_proposals", C.c_int64),
        ("ms_total", C.c_double),
        ("ms_merge", C.c_double),
        ("ms_carry", C.c_double),
        ("ms_nodal", C.c_double),
        ("ms_entropy", C.c_double),
        ("ms_golden", C.c_double),
        ("rebuilds", C.c_int32),
        ("error", C.c_int32),
        ("grows", C.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class UsapPhase(C.Structure):
    _fields_ = [
        ("B", C.c_int32),
        ("sweeps", C.c_int32),
        ("proposals", C.c_int64),
        ("moves", C.c_int64),
        ("S", C.c_double),
        ("ms_merge", C.c_double),
        ("ms_nodal", C.c_double),
        ("k_nodal_ms", C.c_double),
        ("k_apply_ms", C.c_double),
        ("k_merge_ms", C.c_double),
        ("k_build_ms", C.c_double),
        ("k_grow_ms", C.c_double),
        ("rebuilds", C.c_int32),
        ("merge_proposals", C.c_int64),
    ]


class UsapSweep(C.Structure):
    _fields_ = [
        ("batches", C.c_int32),
        ("converged", C.c_int32),
        ("proposals", C.c_int64),
        ("moves", C.c_int64),
        ("S", C.c_double),
        ("dS_sum", C.c_double),
        ("dS_hi", C.c_int64),
        ("dS_lo", C.c_uint64),
        ("dS_inf", C.c_int32),
        ("rebuilds", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class UsapBatchStats(C.Structure):
    _fields_ = [
        ("workers", C.c_int32),
        ("failed", C.c_int32),
        ("ms_total", C.c_double),
        ("graphs_per_s", C.c_double),
        ("nodal_proposals", C.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class UsapServeGraph(C.Structure):
    _fields_ = [
        ("N", C.c_int32),
        ("E", C.c_int64),
        ("src", C.c_void_p),
        ("dst", C.c_void_p),
        ("w", C.c_void_p),
        ("arrival_s", C.c_double),
        ("params", UsapParams),
    ]


class UsapServeOpts(C.Structure):
    _fields_ = [
        ("stages", C.c_int32),
        ("interval_s", C.c_double),
        ("workers", C.c_int32),
        ("hbm_budget", C.c_int64),
        ("final_partition", C.c_int32),
        ("repartition_growth", C.c_double),
    ]


class UsapServeStage(C.Structure):
    _fields_ = [
        ("graph", C.c_int32),
        ("stage", C.c_int32),
        ("t_ready", C.c_double),
        ("t_start", C.c_double),
        ("t_done", C.c_double),
        ("t_admit", C.c_double),
        ("t_release", C.c_double),
        ("S", C.c_double),
        ("B", C.c_int32),
        ("error", C.c_int32),
        ("E", C.c_int64),
        ("sweeps", C.c_int64),
        ("nodal_proposals", C.c_int64),
        ("full", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class UsapServeStats(C.Structure):
    _fields_ = [
        ("wall_s", C.c_double),
        ("graphs_per_s", C.c_double),
        ("p50_ms", C.c_double),
        ("p99_ms", C.c_double),
        ("p50_first_ms", C.c_double),
        ("p99_first_ms", C.c_double),
        ("p50_resume_ms", C.c_double),
        ("p99_resume_ms", C.c_double),
        ("max_resident", C.c_int32),
        ("failed", C.c_int32),
        ("max_bytes", C.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class UsapKernelProf(C.Structure):
    _fields_ = [("launches", C.c_int64), ("ms", C.c_double), ("bytes", C.c_double),
                ("model_bytes", C.c_double)]


KERNELS = ["nodal", "apply", "merge", "entropy", "build", "grow"]


# (name, restype, argtypes) of every exported entry point of include/usap.h
SIGNATURES = [
    ("usap_default_params", None, [C.POINTER(UsapParams)]),
    ("usap_create", C.c_int, [C.POINTER(UsapGraph), C.POINTER(UsapParams), C.POINTER(_vp)]),
    ("usap_destroy", C.c_int, [_vp]),
    ("usap_last_error", C.c_char_p, [_vp]),
    ("usap_set_graph", C.c_int, [_vp, C.POINTER(UsapGraph)]),
    ("usap_append_edges", C.c_int, [_vp, C.c_int64, _vp, _vp, _vp]),
    ("usap_create_from_edges", C.c_int, [C.c_int32, C.c_int64, _vp, _vp, _vp,
                                         C.POINTER(UsapParams), C.POINTER(_vp)]),
    ("usap_resume", C.c_int, [_vp, C.c_uint64, _vp, C.POINTER(UsapStats)]),
    ("usap_set_labels", C.c_int, [_vp, _i32p, C.c_int32]),
    ("usap_get_labels", C.c_int, [_vp, _i32p]),
    ("usap_num_blocks", C.c_int, [_vp, C.POINTER(C.c_int32)]),
    ("usap_get_degrees", C.c_int, [_vp, _i32p, _i32p]),
    ("usap_get_M_coo", C.c_int, [_vp, C.c_int64, _vp, _vp, _vp, C.POINTER(C.c_int64)]),
    ("usap_entropy", C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                               C.POINTER(C.c_uint64)]),
    ("usap_entropy_noalloc", C.c_int, [_vp, _vp, _vp]),
    ("usap_set_uniforms", C.c_int, [_vp, _vp, C.c_size_t]),
    ("usap_uniforms_used", C.c_int, [_vp, _vp]),
    ("usap_merge_propose", C.c_int, [_vp, C.c_int32, C.c_uint64, _vp, _vp]),
    ("usap_merge_result_fx", C.c_int, [_vp, _i32p, _i64p, _u64p]),
    ("usap_merge_apply", C.c_int, [_vp, C.c_int32, C.POINTER(C.c_int32)]),
    ("usap_nodal_batch", C.c_int, [_vp, C.c_int32, C.c_int32, C.c_uint64,
                                   C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    ("usap_nodal_sweep", C.c_int, [_vp, C.c_int32, C.c_uint64, C.c_double, C.c_double,
                                   C.c_double, C.c_int32, C.POINTER(UsapSweep)]),
    ("usap_nodal_sum_fx", C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_int32)]),
    ("usap_nodal_detail", C.c_int, [_vp, _i32p, _i64p, _u64p, _i32p, _i32p]),
    ("usap_snapshot", C.c_int, [_vp, C.c_int32, C.c_double]),
    ("usap_restore", C.c_int, [_vp, C.c_int32]),
    ("usap_partition", C.c_int, [_vp, _vp, C.POINTER(UsapStats)]),
    ("usap_partition_log", C.c_int, [_vp, C.c_int32, _vp, C.POINTER(C.c_int32)]),
    ("usap_profile", C.c_int, [_vp, C.c_int32, _vp]),
    ("usap_partition_batch", C.c_int, [C.c_int32, _vp, _vp, _vp, _vp, _vp, C.c_int32,
                                       C.POINTER(UsapBatchStats)]),
    ("usap_batch_autotune", C.c_int, [C.c_int32, _vp, _vp, C.c_int32, C.c_int32,
                                      C.POINTER(C.c_int32), _vp, _vp]),
    ("usap_footprint", C.c_int64, [C.c_int32, C.c_int64]),
    ("usap_serve", C.c_int, [C.c_int32, _vp, C.POINTER(UsapServeOpts), _vp, _vp, _vp,
                             C.POINTER(UsapServeStats)]),
    ("usap_golden_update", C.c_int, [_i64p, _f64p, _i64p, C.c_int64, C.c_double, C.c_int64,
                                     C.c_double, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                     C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    ("usap_scc_init", C.c_int, [C.POINTER(UsapGraph), C.c_int32, _i32p, C.POINTER(C.c_int32)]),
    ("usap_model_S", C.c_double, [C.c_int64, C.c_int64, C.c_int64]),
    ("usap_lambda_table", C.c_int, [C.c_int64, _i64p]),
    ("usap_philox4x64", None, [_u64p, _u64p, _u64p]),
    ("usap_build_csr", C.c_int, [C.c_int32, C.c_int64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                 _vp]),
    ("usap_load_tsv", C.c_int, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64), _vp,
                                _vp, _vp]),
    ("usap_generate_dcsbm", C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_int32,
                                      C.c_int32, C.c_double, C.c_double, C.c_uint64,
                                      C.POINTER(C.c_int64), _vp, _vp, _vp]),
]

_lib = None


def lib():
    """Loads libusap.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build it with `python -m usap_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, eng=None):
    if rc != USAP_OK:
        msg = lib().usap_last_error(eng).decode() if eng else ""
        raise UsapError(rc, msg)
    return rc


def ptr(a):
    return None if a is None else a.ctypes.data
