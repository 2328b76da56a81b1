// driver.cpp -- host C++ restatement of Graph_P::partition()
// (graph_parallel.hpp:322-512) and its golden-section search over B
// (graph_parallel.hpp:1322-1365), running the hot path through the engine's
// C-ABI (usap_merge_propose / usap_merge_apply / usap_nodal_sweep /
// usap_entropy / usap_snapshot / usap_restore).
//
// Random streams: merge phase p draws with stream p, the j-th nodal batch of
// the whole run with stream j.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <vector>

#include "engine_internal.hpp"
#include "usap.h"
#include "usap_common.hpp"

namespace {

struct Snap {
  int64_t B = 0;
  double S = INFINITY;
  int64_t id = -1;
};

// _prepare_for_partition_next, graph_parallel.hpp:1322-1365.  Returns 1 when
// the optimal B is found, 0 to continue.  When the entropy already rises on
// the second halving, small is set while large never was; the reference's
// size_t arithmetic is undefined there (it restores an empty partition), so
// the search stops and returns med, the best partition found.
int golden(Snap& large, Snap& med, Snap& small, const Snap& cur, double rate, Snap& next,
           int64_t& B_to_merge) {
  if (cur.S <= med.S) {
    if (med.B > cur.B)
      large = med;
    else
      small = med;
    med = cur;
  } else {
    if (med.B > cur.B)
      small = cur;
    else
      large = cur;
  }
  B_to_merge = 0;
  if (std::isinf(small.S)) {
    next = med;
    B_to_merge = (int64_t)std::floor((double)med.B * rate);
    return B_to_merge == 0 ? 1 : 0;
  }
  if (large.id < 0 || large.B - small.B == 2) {
    next = med;
    return 1;
  }
  if ((large.B - med.B) >= (med.B - small.B)) {
    const int64_t nb = med.B + (int64_t)std::round((double)(large.B - med.B) * 0.618);
    next = large;
    B_to_merge = large.B - nb;
  } else {
    const int64_t nb = small.B + (int64_t)std::round((double)(med.B - small.B) * 0.618);
    next = med;
    B_to_merge = med.B - nb;
  }
  return 0;
}

using clk = std::chrono::steady_clock;
double ms_since(clk::time_point t0) {
  return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

// The nodal phase at a fixed B (graph_parallel.hpp:381-478): sweeps of
// batches until the moving-average convergence test fires (threshold thr) or
// max_nodal_itr sweeps.  The reference recomputes the entropy before each
// sweep (prepare_batch); S only enters the convergence test, so it is
// computed for the sweeps that run the test (itr >= avg_window - 1) and the
// sweeps before run with S = 0 -- the moves, the dS sums and the stop
// decisions are the same.  Advances nodal_stream by the batches run; S
// receives the last sweep's S (callers recompute the MDL after the phase).
int nodal_phase(usap_engine* e, const usap_params& p, int32_t batch, int32_t B, double thr,
                uint64_t& nodal_stream, std::vector<double>& itr_dS, usap_stats& st, double& S) {
  const int32_t W = p.avg_window;
  const int64_t moves0 = st.accepted_moves;
  int rc = USAP_OK;
  std::fill(itr_dS.begin(), itr_dS.end(), 0.0);
  for (int32_t itr = 0; itr < p.max_nodal_itr; ++itr) {
    // prepare_batch: the sweep's starting S (the entropy of the current
    // state, only needed by the convergence test) is computed on the device
    // ahead of the sweep's first batch
    const bool test = itr >= W - 1;
    const double no_S = 0.0;
    double mean = 0;
    if (itr >= W - 1) {
      for (int32_t i = itr - W + 1; i < itr; ++i) mean += itr_dS[i];
      mean /= (double)(W - 1);
    }
    st.sweeps++;
    const auto t0 = clk::now();
    // fetch_batch / nodal_move / update / check_batch (graph_parallel.hpp:397-468),
    // the update step and its convergence test run on the device
    usap_sweep sw{};
    if ((rc = usap::nodal_sweep(e, batch, nodal_stream, test ? nullptr : &no_S, mean, thr, test,
                                &sw)))
      return rc;
    nodal_stream += (uint64_t)sw.batches;
    st.batches += sw.batches;
    st.nodal_proposals += sw.proposals;
    st.accepted_moves += sw.moves;
    itr_dS[itr] = sw.dS_sum;
    S = sw.S;
    st.ms_nodal += ms_since(t0);
    if (p.verbose > 1)
      std::fprintf(stderr, "usap:   sweep %d B=%d S=%.3f moves=%lld %.1f ms\n", itr, B, S,
                   (long long)(st.accepted_moves - moves0), ms_since(t0));
    if (sw.converged) break;
  }
  return rc;
}

}  // namespace

extern "C" {

int usap_golden_update(int64_t* slot_B, double* slot_S, int64_t* slot_id, int64_t cur_B,
                       double cur_S, int64_t cur_id, double rate, int64_t* next_B,
                       double* next_S, int64_t* next_id, int64_t* B_to_merge) {
  if (!slot_B || !slot_S || !slot_id || !next_B || !next_S || !next_id || !B_to_merge)
    return USAP_EINVAL;
  Snap s[3];
  for (int i = 0; i < 3; ++i) s[i] = Snap{slot_B[i], slot_S[i], slot_id[i]};
  Snap cur{cur_B, cur_S, cur_id}, next;
  int64_t btm = 0;
  const int rc = golden(s[0], s[1], s[2], cur, rate, next, btm);
  for (int i = 0; i < 3; ++i) {
    slot_B[i] = s[i].B;
    slot_S[i] = s[i].S;
    slot_id[i] = s[i].id;
  }
  *next_B = next.B;
  *next_S = next.S;
  *next_id = next.id;
  *B_to_merge = btm;
  return rc;
}

int usap_partition(usap_engine* e, int32_t* labels_out, usap_stats* stats) {
  if (!e) return USAP_EINVAL;
  const usap_graph* g = usap::engine_host_graph(e);
  if (!g) return USAP_EHIP;
  const usap_params& p = *usap::engine_params(e);
  const int32_t N = g->N;
  usap_stats st{};
  const int rebuilds0 = usap::engine_rebuilds(e);
  const int64_t grows0 = usap::engine_grows(e);
  const auto t_all = clk::now();
  int rc = USAP_OK;

  // init_merge (graph_parallel.hpp:338-342)
  std::vector<int32_t> lab((size_t)N);
  int32_t B = 0;
  if (p.init == USAP_INIT_IDENTITY) {
    std::iota(lab.begin(), lab.end(), 0);
    B = N;
  } else {
    if ((rc = usap_scc_init(g, p.dfs_depth, lab.data(), &B))) return rc;
  }
  if ((rc = usap_set_labels(e, lab.data(), B))) return rc;
  int64_t B_to_merge = (int64_t)std::floor((double)B * p.reduction_rate);

  int32_t batch = p.nodal_batch > 0 ? p.nodal_batch : N / p.num_batch_nodal_update;
  if (batch < 1) batch = 1;
  std::vector<double> itr_dS((size_t)p.max_nodal_itr + 1, 0.0);
  Snap large, med, small;
  // snapshot slot of each partition id held by the golden-section bracket
  std::vector<int> slot_of;
  bool slot_used[4] = {false, false, false, false};
  uint64_t merge_stream = 0, nodal_stream = 0;
  double S = 0;
  std::vector<usap_phase>& plog = usap::engine_phase_log(e);
  plog.clear();

  for (int64_t phase = 0;; ++phase) {
    // prepare_merge + block_merge + perform_merge + prepare_move
    usap_phase ph{};
    const int64_t sweeps0 = st.sweeps, props0 = st.nodal_proposals, moves0 = st.accepted_moves;
    usap_kernel_prof kp0[USAP_NKERNELS];
    usap_profile(e, -1, kp0);
    const int32_t rb0 = usap::engine_rebuilds(e);
    const auto t_phase = clk::now();
    auto t0 = clk::now();
    if ((rc = usap_merge_propose(e, p.K, merge_stream++, nullptr, nullptr))) break;
    // wait for the proposals here so that ms_merge holds the merge kernel
    if ((rc = usap_merge_result_fx(e, nullptr, nullptr, nullptr))) break;
    st.merge_proposals += (int64_t)B * p.K;
    ph.merge_proposals = (int64_t)B * p.K;
    st.ms_merge += ms_since(t0);
    t0 = clk::now();
    if ((rc = usap_merge_apply(e, (int32_t)B_to_merge, &B))) break;
    st.ms_carry += ms_since(t0);
    st.phases++;
    ph.B = B;
    ph.ms_merge = ms_since(t_phase);
    const auto t_nodal = clk::now();
    const bool isf = std::isfinite(large.S) && std::isfinite(med.S) && std::isfinite(small.S);
    rc = nodal_phase(e, p, batch, B, isf ? p.thr2 : p.thr1, nodal_stream, itr_dS, st, S);
    if (rc) break;
    ph.ms_nodal = ms_since(t_nodal);
    // prepare_for_next
    t0 = clk::now();
    if ((rc = usap_entropy(e, &S, nullptr, nullptr))) break;
    st.ms_entropy += ms_since(t0);
    ph.sweeps = (int32_t)(st.sweeps - sweeps0);
    ph.proposals = st.nodal_proposals - props0;
    ph.moves = st.accepted_moves - moves0;
    ph.S = S;
    usap_kernel_prof kp1[USAP_NKERNELS];
    usap_profile(e, -1, kp1);
    ph.k_nodal_ms = kp1[USAP_K_NODAL].ms - kp0[USAP_K_NODAL].ms;
    ph.k_apply_ms = kp1[USAP_K_APPLY].ms - kp0[USAP_K_APPLY].ms;
    ph.k_merge_ms = kp1[USAP_K_MERGE].ms - kp0[USAP_K_MERGE].ms;
    ph.k_build_ms = kp1[USAP_K_BUILD].ms - kp0[USAP_K_BUILD].ms;
    ph.k_grow_ms = kp1[USAP_K_GROW].ms - kp0[USAP_K_GROW].ms;
    ph.rebuilds = usap::engine_rebuilds(e) - rb0;
    plog.push_back(ph);
    t0 = clk::now();
    int slot = -1;
    for (int i = 0; i < 4; ++i)
      if (!slot_used[i]) {
        slot = i;
        break;
      }
    if (slot < 0) {
      rc = USAP_ESTATE;
      break;
    }
    if ((rc = usap_snapshot(e, slot, S))) break;
    slot_used[slot] = true;
    slot_of.push_back(slot);
    Snap cur{B, S, phase}, next;
    const int gr = golden(large, med, small, cur, p.reduction_rate, next, B_to_merge);
    if (next.id != phase) {
      if ((rc = usap_restore(e, slot_of[next.id]))) break;
      B = (int32_t)next.B;
    }
    for (int64_t id = 0; id <= phase; ++id) {
      const int sl = slot_of[id];
      if (sl < 0) continue;
      if (id == large.id || id == med.id || id == small.id || id == next.id) continue;
      slot_used[sl] = false;
      slot_of[id] = -1;
    }
    st.ms_golden += ms_since(t0);
    if (p.verbose) {
      long long gd[4];
      usap::engine_grow_debug(e, gd);
      std::fprintf(stderr, "usap:   growth: grown %lld compacted %lld failed-inserts %lld pool-full %lld\n",
                   gd[0], gd[1], gd[2], gd[3]);
    }
    if (p.verbose)
      std::fprintf(stderr,
                   "usap: phase %lld B=%d S=%.6f bracket=[%lld,%lld,%lld] next B=%lld sweeps=%d "
                   "ms merge=%.1f nodal=%.1f k_nodal=%.1f k_apply=%.1f k_grow=%.1f build=%.1f rebuilds=%d\n",
                   (long long)phase, ph.B, ph.S, (long long)large.B, (long long)med.B,
                   (long long)small.B, (long long)(next.B - B_to_merge), ph.sweeps, ph.ms_merge,
                   ph.ms_nodal, ph.k_nodal_ms, ph.k_apply_ms, ph.k_grow_ms, ph.k_build_ms, ph.rebuilds);
    if (gr == 1) {
      S = next.S;
      break;
    }
  }
  st.ms_total = ms_since(t_all);
  st.S = S;
  st.B = B;
  st.rebuilds = usap::engine_rebuilds(e) - rebuilds0;
  st.grows = usap::engine_grows(e) - grows0;
  st.error = rc;
  if (stats) *stats = st;
  if (rc) return rc;
  if (labels_out) rc = usap_get_labels(e, labels_out);
  return rc;
}

int usap_resume(usap_engine* e, uint64_t stream0, int32_t* labels_out, usap_stats* stats) {
  if (!e) return USAP_EINVAL;
  const usap_graph* g = usap::engine_host_graph(e);
  if (!g) return USAP_EHIP;
  const usap_params& p = *usap::engine_params(e);
  usap_stats st{};
  int32_t B = 0;
  int rc = usap_num_blocks(e, &B);
  if (rc) return rc;
  const int rebuilds0 = usap::engine_rebuilds(e);
  const int64_t grows0 = usap::engine_grows(e);
  const auto t_all = clk::now();
  int32_t batch = p.nodal_batch > 0 ? p.nodal_batch : g->N / p.num_batch_nodal_update;
  if (batch < 1) batch = 1;
  std::vector<double> itr_dS((size_t)p.max_nodal_itr + 1, 0.0);
  uint64_t nodal_stream = stream0;
  double S = 0;
  rc = nodal_phase(e, p, batch, B, p.thr2, nodal_stream, itr_dS, st, S);
  if (rc == USAP_OK) {
    const auto t0 = clk::now();
    rc = usap_entropy(e, &S, nullptr, nullptr);
    st.ms_entropy += ms_since(t0);
  }
  st.ms_total = ms_since(t_all);
  st.S = S;
  st.B = B;
  st.rebuilds = usap::engine_rebuilds(e) - rebuilds0;
  st.grows = usap::engine_grows(e) - grows0;
  st.error = rc;
  if (stats) *stats = st;
  if (rc) return rc;
  if (labels_out) rc = usap_get_labels(e, labels_out);
  return rc;
}

}  // extern "C"
