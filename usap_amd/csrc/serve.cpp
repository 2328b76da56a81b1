// serve.cpp -- a serving loop for graphs that arrive over time as streamed
// edge increments.
//
// No counterpart in the reference (it lists streaming as future work,
// README.md:47-49).  Graph i arrives at arrival_s; its edges come in `stages`
// increments, interval_s apart: stage k holds the first ceil(E*k/stages)
// edges of its delivery order.  Stage 1 builds the engine from its edges on
// the device (usap_create_from_edges) and runs a full partition()
// (usap_partition); stage k > 1 appends its increment on the device
// (usap_append_edges, labels kept) and resumes the nodal chain from the
// previous stage's partition (usap_resume).  The resumed chain keeps the
// block count of the last full partition, and merges cannot add blocks, so
// a stage runs a full partition again once the graph has grown by
// `repartition_growth` since the last one, and, optionally, at the last
// stage.  A graph's engine -- lists, labels, M, degrees -- stays resident on
// the device from its first stage to its last.
//
// Continuous batching: `workers` host threads take whatever stage is ready
// (its increment has arrived and the graph's previous stage is done), oldest
// first, so stages of different graphs run concurrently on the device, each
// on its own engine stream.  A graph is admitted (its engine may be created)
// in arrival order once its estimated device footprint fits the budget next
// to the resident graphs; it leaves after its last stage.  Every result is
// the one the graph gets alone: engines are deterministic and independent.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine_internal.hpp"
#include "usap.h"

namespace {

using clk = std::chrono::steady_clock;

struct GraphState {
  usap_engine* eng = nullptr;
  int32_t next = 1;        // next stage to run (1-based)
  bool admitted = false;
  bool running = false;
  bool done = false;
  int64_t bytes = 0;       // footprint charged to the budget while resident
  int64_t E_full = 0;      // edges at the graph's last full partition
};

struct Server {
  int32_t n = 0;
  const usap_serve_graph* graphs = nullptr;
  usap_serve_opts opt{};
  int32_t* labels = nullptr;
  const int64_t* label_off = nullptr;
  usap_serve_stage* stages = nullptr;

  std::mutex mu;
  std::condition_variable cv;
  clk::time_point t0;
  std::vector<GraphState> gs;
  std::vector<int32_t> by_arrival;  // graph ids by arrival time
  size_t arrived = 0;               // graphs of by_arrival that have arrived
  std::deque<int32_t> waiting;      // arrived, not yet admitted
  int64_t resident_bytes = 0, max_bytes = 0;
  int32_t resident = 0, max_resident = 0, finished = 0, failed = 0;

  double now() const { return std::chrono::duration<double>(clk::now() - t0).count(); }
  double ready_at(int32_t g, int32_t k) const {
    return graphs[g].arrival_s + (double)(k - 1) * opt.interval_s;
  }

  // Under mu: moves arrivals into the queue and admits from its head.
  void admit(double t) {
    while (arrived < by_arrival.size() && graphs[by_arrival[arrived]].arrival_s <= t)
      waiting.push_back(by_arrival[arrived++]);
    while (!waiting.empty()) {
      const int32_t g = waiting.front();
      const int64_t b = usap_footprint(graphs[g].N, graphs[g].E);
      if (opt.hbm_budget > 0 && resident > 0 && resident_bytes + b > opt.hbm_budget) break;
      waiting.pop_front();
      gs[g].admitted = true;
      gs[g].bytes = b;
      resident_bytes += b;
      ++resident;
      max_bytes = std::max(max_bytes, resident_bytes);
      max_resident = std::max(max_resident, resident);
      stages[(int64_t)g * opt.stages].t_admit = t;
    }
  }

  // Under mu: the ready stage that has waited longest, or -1.
  int32_t pick(double t) const {
    int32_t best = -1;
    double best_t = 0;
    for (int32_t g = 0; g < n; ++g) {
      const GraphState& s = gs[g];
      if (!s.admitted || s.running || s.done) continue;
      const double r = ready_at(g, s.next);
      if (r > t) continue;
      if (best < 0 || r < best_t) {
        best = g;
        best_t = r;
      }
    }
    return best;
  }

  // Under mu: when something may next become ready (arrival or increment).
  double next_event() const {
    double t = INFINITY;
    if (arrived < by_arrival.size()) t = graphs[by_arrival[arrived]].arrival_s;
    for (int32_t g = 0; g < n; ++g) {
      const GraphState& s = gs[g];
      if (s.admitted && !s.running && !s.done) t = std::min(t, ready_at(g, s.next));
    }
    return t;
  }

  // Runs stage k of graph g (outside mu).  Only the stage's increment goes
  // to the device: stage 1 builds the engine's lists from its edges, later
  // stages append theirs.
  int run_stage(int32_t g, int32_t k, usap_serve_stage& rec) {
    const usap_serve_graph& G = graphs[g];
    const int64_t Ek = (int64_t)std::ceil((double)G.E * k / opt.stages);
    const int64_t Ep = (int64_t)std::ceil((double)G.E * (k - 1) / opt.stages);
    usap_stats st;
    std::memset(&st, 0, sizeof(st));
    usap_engine*& e = gs[g].eng;
    int32_t* lab = labels + label_off[g];
    int rc = USAP_OK;
    if (k == 1) {
      rc = usap_create_from_edges(G.N, Ek, G.src, G.dst, G.w, &G.params, &e);
      if (rc == USAP_OK) rc = usap_partition(e, lab, &st);
      gs[g].E_full = Ek;
      rec.full = 1;
    } else {
      rc = usap_append_edges(e, Ek - Ep, G.src + Ep, G.dst + Ep, G.w ? G.w + Ep : nullptr);
      // a full partition (merge phases and the search over B) at the last
      // stage, or once the graph has grown by the given factor since the
      // last one; else the nodal chain resumes at the current B
      const bool full = (k == opt.stages && opt.final_partition) ||
                        (opt.repartition_growth > 1.0 &&
                         (double)Ek >= opt.repartition_growth * (double)gs[g].E_full);
      if (rc == USAP_OK && full) {
        rc = usap_partition(e, lab, &st);
        gs[g].E_full = Ek;
        rec.full = 1;
      } else if (rc == USAP_OK) {
        // each stage draws from its own range of nodal streams
        rc = usap_resume(e, (uint64_t)k << 40, lab, &st);
      }
    }
    rec.S = st.S;
    rec.B = st.B;
    rec.sweeps = st.sweeps;
    rec.nodal_proposals = st.nodal_proposals;
    rec.E = Ek;
    return rc;
  }

  void worker() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      if (finished == n) return;
      const double t = now();
      admit(t);
      const int32_t g = pick(t);
      if (g < 0) {
        const double te = next_event();
        if (std::isfinite(te))
          cv.wait_for(lk, std::chrono::duration<double>(std::max(0.0, te - t) + 1e-4));
        else
          cv.wait(lk);
        continue;
      }
      GraphState& s = gs[g];
      const int32_t k = s.next;
      s.running = true;
      usap_serve_stage& rec = stages[(int64_t)g * opt.stages + (k - 1)];
      rec.graph = g;
      rec.stage = k;
      rec.t_ready = ready_at(g, k);
      rec.t_start = t;
      lk.unlock();
      int rc = run_stage(g, k, rec);
      if (rc != USAP_OK)
        usap::set_detached_error("usap_serve: graph " + std::to_string(g) + " stage " +
                                 std::to_string(k) + ": error " + std::to_string(rc) + ": " +
                                 (s.eng ? usap_last_error(s.eng) : "usap_create failed"));
      lk.lock();
      rec.t_done = now();
      rec.error = rc;
      s.running = false;
      s.next = k + 1;
      if (rc != USAP_OK || k == opt.stages) {  // leaves: free its device memory
        if (s.eng) usap_destroy(s.eng);
        s.eng = nullptr;
        s.done = true;
        if (rc != USAP_OK) ++failed;
        resident_bytes -= s.bytes;
        --resident;
        ++finished;
        stages[(int64_t)g * opt.stages].t_release = rec.t_done;
      }
      cv.notify_all();
    }
  }
};

double pct(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const size_t i = (size_t)std::min<double>((double)v.size() - 1,
                                            std::ceil(q * (double)v.size()) - 1);
  return v[i];
}

}  // namespace

extern "C" {

int64_t usap_footprint(int32_t N, int64_t E) {
  if (N < 1 || E < 0) return 0;
  const int64_t n = N, m = E;
  const int64_t graph = 16 * (n + 1) + 20 * m + 8 * (m + 1) + 8 * n + 96 * m;  // + lambda, large-node tables
  const int64_t state = 160 * n + 24 * std::min<int64_t>(1 << 22, 2 * m + 4096) + 65536;
  const int64_t buckets = n * 512 * 4;  // bucket sums at B = N
  // block-row slabs: 16-byte slots, three times the rows' capacities (each a
  // power of two below 2 * (2 * (out + in edges) + 16), slab_slack 2) plus
  // headroom
  const int64_t pools = 16 * (3 * (8 * m + 32 * n) + 4 * n + 4096);
  return graph + state + buckets + pools;
}

int usap_serve(int32_t n, const usap_serve_graph* graphs, const usap_serve_opts* opt,
               int32_t* labels, const int64_t* label_off, usap_serve_stage* stages,
               usap_serve_stats* out) {
  if (n < 0 || !opt || opt->stages < 1 || opt->workers < 1 || opt->interval_s < 0 ||
      opt->hbm_budget < 0)
    return USAP_EINVAL;
  if (n > 0 && (!graphs || !labels || !label_off || !stages)) return USAP_EINVAL;
  for (int32_t i = 0; i < n; ++i)
    if (graphs[i].N < 1 || graphs[i].E < 0 || (graphs[i].E > 0 && (!graphs[i].src || !graphs[i].dst)) ||
        graphs[i].arrival_s < 0)
      return USAP_EINVAL;
  usap::clear_detached_error();
  // captured-sweep replay only with one stage in flight, as in
  // usap_partition_batch
  if (opt->workers > 1 && n > 1)
    for (int32_t i = 0; i < n; ++i)
      if (graphs[i].params.graphs) {
        usap::set_detached_error("usap_serve: graphs=1 (captured-sweep replay) needs workers=1");
        return USAP_EINVAL;
      }
  Server sv;
  sv.n = n;
  sv.graphs = graphs;
  sv.opt = *opt;
  sv.labels = labels;
  sv.label_off = label_off;
  sv.stages = stages;
  sv.gs.resize(n);
  std::memset(stages, 0, sizeof(usap_serve_stage) * (size_t)n * opt->stages);
  sv.by_arrival.resize(n);
  for (int32_t i = 0; i < n; ++i) sv.by_arrival[i] = i;
  std::stable_sort(sv.by_arrival.begin(), sv.by_arrival.end(),
                   [&](int32_t a, int32_t b) { return graphs[a].arrival_s < graphs[b].arrival_s; });
  sv.t0 = clk::now();
  std::vector<std::thread> pool;
  for (int32_t w = 1; w < opt->workers; ++w) {
    try {
      pool.emplace_back([&sv]() { sv.worker(); });
    } catch (...) {
      break;
    }
  }
  sv.worker();
  for (std::thread& t : pool) t.join();
  const double wall = sv.now();
  if (out) {
    std::memset(out, 0, sizeof(*out));
    std::vector<double> all, first, later;
    for (int64_t i = 0; i < (int64_t)n * opt->stages; ++i) {
      const usap_serve_stage& r = stages[i];
      if (r.stage < 1) continue;
      const double ms = 1e3 * (r.t_done - r.t_ready);
      all.push_back(ms);
      (r.stage == 1 ? first : later).push_back(ms);
    }
    double first_arrival = INFINITY, last_done = 0;
    for (int32_t i = 0; i < n; ++i) first_arrival = std::min(first_arrival, graphs[i].arrival_s);
    for (int64_t i = 0; i < (int64_t)n * opt->stages; ++i) last_done = std::max(last_done, stages[i].t_done);
    out->wall_s = wall;
    out->graphs_per_s = n > 0 && last_done > first_arrival ? n / (last_done - first_arrival) : 0.0;
    out->p50_ms = pct(all, 0.50);
    out->p99_ms = pct(all, 0.99);
    out->p50_first_ms = pct(first, 0.50);
    out->p99_first_ms = pct(first, 0.99);
    out->p50_resume_ms = pct(later, 0.50);
    out->p99_resume_ms = pct(later, 0.99);
    out->max_resident = sv.max_resident;
    out->max_bytes = sv.max_bytes;
    out->failed = sv.failed;
  }
  return sv.failed ? USAP_ESTATE : USAP_OK;
}

}  // extern "C"
