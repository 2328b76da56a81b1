// usap/graph_p.hpp -- drop-in C++ front end with the public interface of
// sgp::Graph_P<W> (chih-chun-chang/uSAP include/graph_parallel.hpp:27-245),
// running partition() on the MI355X engine through the C-ABI of usap.h.
//
//   usap::Graph_P<long> g("path/to/graph");   // reads graph.tsv + graph_truePartition.tsv
//   g.partition();
//   const std::vector<size_t>& p = g.get_partitions();
//
// Link with usap_amd/lib/libusap.so.  Errors throw std::runtime_error (the
// reference exits on unreadable files, graph_parallel.hpp:254-256).
#pragma once

#include <cstdint>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../usap.h"

namespace usap {

template <typename W = long>
class Graph_P {
 public:
  // graph_parallel.hpp:41-52 (beta, verbose and t_block_num are unused there too)
  int beta = 3;
  size_t block_size = 1024;
  size_t num_agg_proposals_per_block = 10;
  float num_block_reduction_rate = 0.5;
  size_t max_num_nodal_itr = 100;
  size_t num_batch_nodal_update = 4;
  float delta_entropy_threshold1 = 5e-4;
  float delta_entropy_threshold2 = 1e-4;
  size_t delta_entropy_moving_avg_window = 3;
  bool verbose = false;
  size_t dfs_depth = 20;
  size_t t_block_num = 20;
  // engine options
  size_t nodal_batch_size = 0;     // 0: N / num_batch_nodal_update
  int init = USAP_INIT_SCC;
  int ds_mode = USAP_DS_REF;
  uint64_t seed = 0;
  int device = 0;

  std::vector<size_t> truePartitions;
  usap_stats stats{};

  explicit Graph_P(const std::string& FileName, size_t num_threads = 16)
      : num_threads_(num_threads) {
    load_graph_from_tsv(FileName);
  }

  void load_graph_from_tsv(const std::string& FileName) {
    const std::string path = FileName + ".tsv";
    int32_t N = 0;
    int64_t E = 0;
    if (usap_load_tsv(path.c_str(), &N, &E, nullptr, nullptr, nullptr) != USAP_OK)
      throw std::runtime_error("Unable to open file: " + path);
    std::vector<int32_t> src((size_t)E), dst((size_t)E), w((size_t)E);
    if (usap_load_tsv(path.c_str(), &N, &E, src.data(), dst.data(), w.data()) != USAP_OK)
      throw std::runtime_error("Unable to parse file: " + path);
    bool unit = true;
    for (int32_t x : w) unit = unit && x == 1;
    N_ = (size_t)N;
    E_ = (size_t)E;
    out_ptr_.assign(N_ + 1, 0);
    in_ptr_.assign(N_ + 1, 0);
    out_idx_.assign(E_, 0);
    in_idx_.assign(E_, 0);
    out_w_.clear();
    in_w_.clear();
    if (!unit) {
      out_w_.assign(E_, 0);
      in_w_.assign(E_, 0);
    }
    if (usap_build_csr(N, E, src.data(), dst.data(), unit ? nullptr : w.data(), out_ptr_.data(),
                       out_idx_.data(), unit ? nullptr : out_w_.data(), in_ptr_.data(),
                       in_idx_.data(), unit ? nullptr : in_w_.data()) != USAP_OK)
      throw std::runtime_error("bad graph: " + path);
    truePartitions.clear();
    std::ifstream tf(FileName + "_truePartition.tsv");
    if (tf.is_open()) {
      truePartitions.assign(N_, 0);
      std::string line;
      while (std::getline(tf, line)) {
        std::istringstream is(line);
        size_t node = 0, block = 0;
        if (is >> node >> block && node >= 1 && node <= N_) truePartitions[node - 1] = block - 1;
      }
    }
  }

  const size_t& num_nodes() const { return N_; }
  const size_t& num_edges() const { return E_; }
  const std::vector<size_t>& get_partitions() const { return partitions_; }

  void partition() {
    usap_graph g{};
    g.N = (int32_t)N_;
    g.E = (int64_t)E_;
    g.out_ptr = out_ptr_.data();
    g.out_idx = out_idx_.data();
    g.in_ptr = in_ptr_.data();
    g.in_idx = in_idx_.data();
    g.out_w = out_w_.empty() ? nullptr : out_w_.data();
    g.in_w = in_w_.empty() ? nullptr : in_w_.data();
    usap_params p;
    usap_default_params(&p);
    p.K = (int32_t)num_agg_proposals_per_block;
    p.reduction_rate = num_block_reduction_rate;
    p.max_nodal_itr = (int32_t)max_num_nodal_itr;
    p.num_batch_nodal_update = (int32_t)num_batch_nodal_update;
    p.nodal_batch = (int32_t)nodal_batch_size;
    p.thr1 = delta_entropy_threshold1;
    p.thr2 = delta_entropy_threshold2;
    p.avg_window = (int32_t)delta_entropy_moving_avg_window;
    p.dfs_depth = (int32_t)dfs_depth;
    p.init = init;
    p.ds_mode = ds_mode;
    p.seed = seed;
    p.device = device;
    p.verbose = verbose ? 1 : 0;
    usap_engine* e = nullptr;
    int rc = usap_create(&g, &p, &e);
    if (rc != USAP_OK) throw std::runtime_error("usap_create failed: " + std::to_string(rc));
    std::vector<int32_t> lab(N_);
    rc = usap_partition(e, lab.data(), &stats);
    const std::string msg = usap_last_error(e);
    usap_destroy(e);
    if (rc != USAP_OK)
      throw std::runtime_error("usap_partition failed (" + std::to_string(rc) + "): " + msg);
    partitions_.assign(lab.begin(), lab.end());
  }

 private:
  size_t num_threads_;  // kept for interface parity: the engine is one GPU
  size_t N_ = 0, E_ = 0;
  std::vector<int64_t> out_ptr_, in_ptr_;
  std::vector<int32_t> out_idx_, in_idx_, out_w_, in_w_;
  std::vector<size_t> partitions_;
};

}  // namespace usap
