"""usap_amd: MI355X-native stochastic block partitioner.

The block-merge and nodal-MCMC hot path of uSAP's Graph_P::partition() as
hand-written CDNA4 HIP kernels behind the C-ABI of include/usap.h
(usap_amd/lib/libusap.so), driven by a host C++ partition driver with the
golden-section search over the block count.
"""
from .batch import autotune, batch_graphs, partition_batch
from .engine import Engine, default_params, golden_update, lambda_table, model_S, philox4x64, scc_init
from .graph import CONFIGS, Graph, config_graph, generate_dcsbm, load_tsv
from .metrics import nmi, pairwise_precision_recall
from .partitioner import Graph_P

__all__ = [
    "Engine", "Graph", "Graph_P", "CONFIGS", "autotune", "batch_graphs", "config_graph",
    "default_params", "generate_dcsbm", "partition_batch",
    "golden_update", "lambda_table", "load_tsv", "model_S", "nmi", "pairwise_precision_recall",
    "philox4x64", "scc_init",
]
