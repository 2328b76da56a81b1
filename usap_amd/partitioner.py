"""Graph_P: the reference's public partitioner interface
(sgp::Graph_P<W>, include/graph_parallel.hpp:27-245) over the MI355X engine.

Same names and meaning: the constructor loads a GraphChallenge TSV pair
(FileName + ".tsv", FileName + "_truePartition.tsv"), the public parameter
fields of graph_parallel.hpp:41-52 are attributes, partition() runs the
partition (graph_parallel.hpp:322-512) and get_partitions() returns the
labels.  Extra attributes select this engine's options (seed, init, ds_mode,
nodal_batch_size, device).
"""
import numpy as np

from .engine import INIT_IDENTITY, INIT_SCC, Engine
from .graph import load_tsv


class Graph_P:
    def __init__(self, FileName=None, num_threads=16, graph=None):
        # graph_parallel.hpp:41-52 (beta, verbose and t_block_num are unused there too)
        self.beta = 3
        self.block_size = 1024
        self.num_agg_proposals_per_block = 10
        self.num_block_reduction_rate = 0.5
        self.max_num_nodal_itr = 100
        self.num_batch_nodal_update = 4
        self.delta_entropy_threshold1 = 5e-4
        self.delta_entropy_threshold2 = 1e-4
        self.delta_entropy_moving_avg_window = 3
        self.verbose = False
        self.dfs_depth = 20
        self.t_block_num = 20
        # engine options
        self.nodal_batch_size = 0
        self.init = INIT_SCC
        self.ds_mode = 0
        self.seed = 0
        self.device = 0
        self.num_threads = num_threads  # kept for interface parity; the engine is one GPU
        self.stats = None
        self._partitions = None
        if graph is not None:
            self.graph = graph
        elif FileName is not None:
            self.load_graph_from_tsv(FileName)
        else:
            raise ValueError("Graph_P needs a FileName or a graph")
        self.truePartitions = self.graph.truth

    def load_graph_from_tsv(self, FileName):
        """graph_parallel.hpp:251-317: FileName.tsv + FileName_truePartition.tsv."""
        import os
        tp = FileName + "_truePartition.tsv"
        self.graph = load_tsv(FileName + ".tsv", tp if os.path.exists(tp) else None)
        self.truePartitions = self.graph.truth

    def num_nodes(self):
        return self.graph.N

    def num_edges(self):
        return self.graph.E

    def get_partitions(self):
        return self._partitions

    def _params(self):
        return dict(K=self.num_agg_proposals_per_block,
                    reduction_rate=self.num_block_reduction_rate,
                    max_nodal_itr=self.max_num_nodal_itr,
                    num_batch_nodal_update=self.num_batch_nodal_update,
                    nodal_batch=self.nodal_batch_size,
                    thr1=self.delta_entropy_threshold1, thr2=self.delta_entropy_threshold2,
                    avg_window=self.delta_entropy_moving_avg_window, dfs_depth=self.dfs_depth,
                    init=self.init, ds_mode=self.ds_mode, seed=self.seed, device=self.device,
                    verbose=int(bool(self.verbose)))

    def partition(self):
        with Engine(self.graph, **self._params()) as eng:
            lab, st = eng.partition()
        self._partitions = lab
        self.stats = st
        return lab


__all__ = ["Graph_P", "INIT_SCC", "INIT_IDENTITY"]
