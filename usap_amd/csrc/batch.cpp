// batch.cpp -- many independent graphs partitioned at once on one GPU.
//
// The reference partitions one graph per process (main.cpp:48-63); its
// per-graph state machine is partition() (graph_parallel.hpp:322-512), which
// usap_partition restates.  Here a pool of host workers each drives one
// engine -- its own HIP stream and device state -- through that state
// machine, taking graphs from one shared queue ordered largest first, so the
// GPU runs the kernels of several graphs concurrently while each graph's
// results stay exactly what it gets alone (every engine is deterministic and
// independent of the others).  A worker's engines share the worker's one HIP
// stream and its own stream-ordered memory pool: engines come and go without
// synchronising the device, and no stream or pool is created or destroyed
// per graph.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "engine_internal.hpp"
#include "usap.h"

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Partitions graphs[idx[0..m)] with `workers` concurrent engines.  labels of
// graph i go to labels + label_off[i]; stats[i] gets its statistics (error
// set on failure).  Returns the number of failed graphs.
int32_t run_pool(const std::vector<int32_t>& idx, const usap_graph* graphs,
                 const usap_params* params, int32_t* labels, const int64_t* label_off,
                 usap_stats* stats, int32_t workers) {
  const int32_t m = (int32_t)idx.size();
  std::atomic<int32_t> next{0};
  std::atomic<int32_t> failed{0};
  auto work = [&]() {
    // this worker's stream and memory pool per device, shared by the engines
    // it creates one after another
    std::vector<std::pair<int32_t, usap::WorkerCtx*>> ctxs;
    auto ctx_for = [&](int32_t dev, usap::WorkerCtx** w) {
      for (auto& c : ctxs)
        if (c.first == dev) {
          *w = c.second;
          return USAP_OK;
        }
      const int rc = usap::worker_ctx_create(dev, w);
      if (rc == USAP_OK) ctxs.emplace_back(dev, *w);
      return rc;
    };
    for (;;) {
      const int32_t k = next.fetch_add(1);
      if (k >= m) break;
      const int32_t i = idx[k];
      usap_engine* e = nullptr;
      usap_stats st;
      std::memset(&st, 0, sizeof(st));
      usap::WorkerCtx* w = nullptr;
      int rc = ctx_for(params[i].device, &w);
      if (rc == USAP_OK) rc = usap::create_in(w, &graphs[i], &params[i], &e);
      if (rc == USAP_OK) rc = usap_partition(e, labels + label_off[i], &st);
      if (rc != USAP_OK)
        usap::set_detached_error("usap_partition_batch: graph " + std::to_string(i) + ": error " +
                                 std::to_string(rc) + ": " + (e ? usap_last_error(e) : "usap_create failed"));
      if (e) usap_destroy(e);
      st.error = rc;
      if (stats) stats[i] = st;
      if (rc != USAP_OK) failed.fetch_add(1);
    }
    for (auto& c : ctxs) usap::worker_ctx_destroy(c.second);
  };
  workers = std::max<int32_t>(1, std::min<int32_t>(workers, m));
  std::vector<std::thread> pool;
  for (int32_t w = 1; w < workers; ++w) {
    try {
      pool.emplace_back(work);
    } catch (...) {
      break;  // fewer threads: the remaining ones take the rest of the queue
    }
  }
  work();
  for (std::thread& t : pool) t.join();
  return failed.load();
}

// Graph order for the queue: most edges first (then most nodes, then index),
// so the largest graphs do not start last.
std::vector<int32_t> size_order(int32_t n, const usap_graph* graphs) {
  std::vector<int32_t> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) {
    if (graphs[a].E != graphs[b].E) return graphs[a].E > graphs[b].E;
    return graphs[a].N > graphs[b].N;
  });
  return idx;
}

bool valid(int32_t n, const usap_graph* graphs, const usap_params* params) {
  if (n < 0 || (n > 0 && (!graphs || !params))) return false;
  for (int32_t i = 0; i < n; ++i)
    if (graphs[i].N < 1) return false;
  return true;
}

// Captured-sweep replay (usap_params.graphs) is refused with several graphs
// in flight: a test process running six engines that capture and replay
// their sweeps on six host threads once ended in SIGABRT (no GPU fault, no
// message from the runtime), and no cause was found.  One graph at a time
// keeps it.
bool replay_in_flight(int32_t n, const usap_params* params, int32_t workers) {
  if (workers == 1 || n < 2) return false;
  for (int32_t i = 0; i < n; ++i)
    if (params[i].graphs) return true;
  return false;
}

}  // namespace

extern "C" {

int usap_partition_batch(int32_t n, const usap_graph* graphs, const usap_params* params,
                         int32_t* labels, const int64_t* label_off, usap_stats* stats,
                         int32_t workers, usap_batch_stats* out) {
  if (!valid(n, graphs, params) || (n > 0 && (!labels || !label_off))) return USAP_EINVAL;
  if (workers < 0) return USAP_EINVAL;
  usap::clear_detached_error();
  if (replay_in_flight(n, params, workers)) {
    usap::set_detached_error(
        "usap_partition_batch: graphs=1 (captured-sweep replay) needs workers=1");
    return USAP_EINVAL;
  }
  if (workers == 0) {
    const int rc = usap_batch_autotune(n, graphs, params, 0, 0, &workers, nullptr, nullptr);
    if (rc != USAP_OK) return rc;
  }
  const std::vector<int32_t> idx = size_order(n, graphs);
  const double t0 = now_ms();
  const int32_t failed = run_pool(idx, graphs, params, labels, label_off, stats, workers);
  const double ms = now_ms() - t0;
  if (out) {
    std::memset(out, 0, sizeof(*out));
    out->workers = std::max<int32_t>(1, std::min<int32_t>(workers, std::max<int32_t>(n, 1)));
    out->failed = failed;
    out->ms_total = ms;
    out->graphs_per_s = ms > 0 ? 1e3 * n / ms : 0.0;
    if (stats)
      for (int32_t i = 0; i < n; ++i) out->nodal_proposals += stats[i].nodal_proposals;
  }
  return failed ? USAP_ESTATE : USAP_OK;
}

int usap_batch_autotune(int32_t n, const usap_graph* graphs, const usap_params* params,
                        int32_t max_workers, int32_t sample, int32_t* workers,
                        int32_t* tried, double* rates) {
  if (!valid(n, graphs, params) || !workers || max_workers < 0 || sample < 0)
    return USAP_EINVAL;
  if (max_workers == 0) max_workers = 16;
  if (sample == 0) sample = 24;
  usap::clear_detached_error();
  if (replay_in_flight(n, params, max_workers)) {
    usap::set_detached_error(
        "usap_batch_autotune: graphs=1 (captured-sweep replay) needs max_workers=1");
    return USAP_EINVAL;
  }
  *workers = 1;
  if (n == 0) return USAP_OK;
  // the sample spans the size range: every k-th graph of the size order
  const std::vector<int32_t> all = size_order(n, graphs);
  std::vector<int32_t> idx;
  const int32_t m = std::min(sample, n);
  for (int32_t j = 0; j < m; ++j) idx.push_back(all[(int64_t)j * n / m]);
  std::vector<int64_t> off(n + 1, 0);
  for (int32_t i = 0; i < n; ++i) off[i + 1] = off[i] + graphs[i].N;
  std::vector<int32_t> scratch((size_t)off[n]);
  // one graph at a time (the plain form), then 2, 4, 8, ... in flight; the
  // best rate wins (the curve is flat near its top, so stopping at the first
  // drop would let one noisy trial decide)
  double best = 0.0;
  int32_t best_w = 1, k = 0;
  for (int32_t w = 1; w <= max_workers; w *= 2) {
    const double t0 = now_ms();
    const int32_t failed =
        run_pool(idx, graphs, params, scratch.data(), off.data(), nullptr, w);
    const double ms = now_ms() - t0;
    if (failed) return USAP_ESTATE;
    const double rate = ms > 0 ? 1e3 * m / ms : 0.0;
    if (tried) tried[k] = w;
    if (rates) rates[k] = rate;
    ++k;
    if (rate > best) {
      best = rate;
      best_w = w;
    }
  }
  *workers = best_w;
  return USAP_OK;
}

}  // extern "C"
