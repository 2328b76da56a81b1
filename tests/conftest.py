import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: longer CPU/GPU case")


# Modules whose tests run several engines on host threads at once.  Two kinds
# of host crash (SIGSEGV, SIGABRT) have ended test processes in such tests,
# with no GPU fault and no cause found; running them after every
# single-engine test keeps a crash there from hiding the parity results.
THREADED = ("test_batch.py", "test_serve.py", "test_gpu_workloads.py")
LAST = ("test_zz_no_alloc.py",)


def pytest_collection_modifyitems(config, items):
    def rank(item):
        name = os.path.basename(str(item.fspath))
        return 2 if name in LAST else 1 if name in THREADED else 0
    items[:] = sorted(items, key=rank)  # stable: file order within each rank


def _gpu_available():
    try:
        import torch  # noqa: F401  (only to count devices)
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def usap():
    import usap_amd
    return usap_amd


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    if not os.path.exists(oracle.LIB_PATH):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       capture_output=True)
    return oracle


def random_graph(N, E, seed, weighted=False, self_loops=False, hubs=0, isolated=0):
    """Small irregular graph: random edges, optional weights, self-loops,
    high-degree hubs (beyond the per-wavefront LDS table) and isolated nodes."""
    rng = np.random.default_rng(seed)
    active = N - isolated
    src = rng.integers(0, active, E)
    dst = rng.integers(0, active, E)
    if not self_loops:
        keep = src != dst
        src, dst = src[keep], dst[keep]
    for h in range(hubs):
        k = 450 + 40 * h
        others = rng.integers(0, active, k)
        src = np.concatenate([src, np.full(k, h), others])
        dst = np.concatenate([dst, others, np.full(k, h)])
        if not self_loops:
            keep = src != dst
            src, dst = src[keep], dst[keep]
    w = rng.integers(1, 4, len(src)) if weighted else None
    import usap_amd
    return usap_amd.Graph.from_edges(src, dst, N, w)
