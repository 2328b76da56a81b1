"""usap_entropy_noalloc (the allocation-free MDL: partial sums in a static
device array, results written to caller-owned device memory, no host wait)
against usap_entropy, bit for bit, on two calls in a row -- the second with a
smaller input -- and once more on a second, smaller engine."""
import ctypes as C

import numpy as np
import pytest

import usap_amd
from usap_amd._lib import lib
from usap_amd.engine import fx as fx_of
from conftest import random_graph

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libusap.so itself links (one runtime per process: a
    second copy, such as the one bundled with torch, would not see the GPU
    once this one holds it)."""
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


def _noalloc(eng):
    hip = _hip()
    buf = C.c_void_p()
    assert hip.hipMalloc(C.byref(buf), 24) == 0  # S (8 bytes) + fx (16 bytes)
    try:
        assert hip.hipMemset(buf, 0xFF, 24) == 0
        S_dev, fx_dev = buf.value, buf.value + 8
        assert lib().usap_entropy_noalloc(eng.h, C.c_void_p(S_dev), C.c_void_p(fx_dev)) == 0
        assert hip.hipDeviceSynchronize() == 0  # the engine's stream has reached the outputs
        S = np.zeros(1, np.float64)
        fx = np.zeros(2, np.int64)
        assert hip.hipMemcpy(S.ctypes.data, C.c_void_p(S_dev), 8, 2) == 0  # device to host
        assert hip.hipMemcpy(fx.ctypes.data, C.c_void_p(fx_dev), 16, 2) == 0
    finally:
        hip.hipFree(buf)
    return fx_of(int(fx[0]), int(fx[1]) & ((1 << 64) - 1)), float(S[0])


def test_entropy_noalloc_equals_entropy(usap):
    print("noalloc: engine 1 (N=6000, identity labels, then one merge round)", flush=True)
    g = random_graph(6000, 60000, seed=51, weighted=True, self_loops=True, hubs=1)
    e1 = usap.Engine(g, seed=5, init=1)
    e1.set_labels(np.arange(g.N, dtype=np.int32), g.N)
    assert _noalloc(e1) == e1.entropy()  # B = N: the largest pool
    e1.merge_propose(0)
    e1.merge_apply(g.N // 2)
    assert _noalloc(e1) == e1.entropy()  # second call, smaller B
    print("noalloc: engine 2 (N=700, SCC labels)", flush=True)
    h = random_graph(700, 5000, seed=52)
    e2 = usap.Engine(h, seed=6)
    lab, B = usap.scc_init(h)
    e2.set_labels(lab, B)
    assert _noalloc(e2) == e2.entropy()
    assert lib().usap_entropy_noalloc(e2.h, None, None) == usap_amd._lib.USAP_EINVAL
    e1.close()
    e2.close()
