/*
 * usap_oracle.h -- CPU restatement of uSAP's partition hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the parity checker and the CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product (usap_amd/, libusap.so) never links, loads or
 * calls it.
 *
 * It restates, in plain C and at the reference's own complexity, the
 * algorithm of include/graph_parallel.hpp (chih-chun-chang/uSAP):
 *   - M build + block degrees         graph_parallel.hpp:703-759
 *   - merge proposal + merge dS       graph_parallel.hpp:761-988
 *   - best merge per block            graph_parallel.hpp:349-368
 *   - carry out best merges           graph_parallel.hpp:1225-1279
 *   - nodal proposal + nodal dS       graph_parallel.hpp:990-1222
 *   - batch apply + M rebuild         graph_parallel.hpp:433-458
 *   - overall entropy (MDL)           graph_parallel.hpp:1281-1319
 *   - golden-section search over B    graph_parallel.hpp:1322-1365
 *   - partition() control flow        graph_parallel.hpp:322-512
 *   - SCC initial partition           graph_parallel.hpp:514-603
 *   - "exact" nodal dS variant        graph.cu:1414-1426
 *
 * Arithmetic (defined by this project, shared with the GPU engine):
 *   lambda(x) = llrint(log(x) * 2^48) as int64, lambda(0) = 0.
 *   T(m,x,y)  = m * (lambda(m) - lambda(x) - lambda(y)) as exact int128,
 *               i.e. 2^48 * m*log(m/(x*y)), and 0 if m == 0, +inf if
 *               m > 0 and x*y == 0.
 *   All sums of T are exact integers, so every summation order gives the
 *   same bits.  A dS is {fx (int128, scale 2^-48), inf flag}.
 *
 * Random stream (defined by this project, shared with the GPU engine):
 *   Philox4x64-10 (Salmon et al., SC'11), key = {seed, 0},
 *   counter = {item, proposal, stream, tag}; the four 64-bit outputs are the
 *   slots 0 = neighbour pick, 1 = accept-uniform test, 2 = uniform block,
 *   3 = multinomial pick.  An index in [0, W) is (x * W) >> 64.
 */
#ifndef USAP_ORACLE_H
#define USAP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_LAMBDA_FRAC 48
#define ORC_TAG_MERGE 1u
#define ORC_TAG_NODAL 2u

typedef struct {
  int32_t N;
  int64_t E;
  const int64_t* out_ptr; /* [N+1] */
  const int32_t* out_idx; /* [E]   out-neighbours, file order */
  const int64_t* in_ptr;  /* [N+1] */
  const int32_t* in_idx;  /* [E]   in-neighbours, file order */
  const int32_t* out_w;   /* [E] or NULL (unweighted) */
  const int32_t* in_w;    /* [E] or NULL */
} orc_graph;

typedef struct {
  int32_t K;                  /* num_agg_proposals_per_block (10) */
  double reduction_rate;      /* num_block_reduction_rate (0.5) */
  int32_t max_nodal_itr;      /* 100 */
  int32_t num_batch_nodal_update; /* 4 -> batch = N/4 */
  int32_t nodal_batch;        /* >0 overrides N/num_batch_nodal_update */
  double thr1;                /* delta_entropy_threshold1 5e-4 */
  double thr2;                /* delta_entropy_threshold2 1e-4 */
  int32_t avg_window;         /* delta_entropy_moving_avg_window 3 */
  int32_t dfs_depth;          /* 20 */
  int32_t block_size;         /* dense/list switch 1024 */
  int32_t init;               /* 0 = SCC (reference), 1 = identity */
  int32_t ds_mode;            /* 0 = ref (graph_parallel.hpp), 1 = exact */
  uint64_t seed;
  int32_t threads;            /* worker threads (16 in the reference) */
} orc_params;

typedef struct orc_state orc_state;

void orc_default_params(orc_params* p);

/* Philox4x64-10: out[4] = philox(ctr[4], key[2]). */
void orc_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);

/* lambda(x) for 0 <= x <= maxv into out[maxv+1]. */
void orc_lambda_table(int64_t maxv, int64_t* out);

/* Inverse-CDF pick: first i with w[0]+..+w[i] > t (t < sum w). */
int64_t orc_pick(const int64_t* w, int64_t n, int64_t t);

/* SCC initial partition (graph_parallel.hpp:514-603). Returns B. */
int32_t orc_scc_init(const orc_graph* g, int32_t dfs_depth, int32_t* labels);

orc_state* orc_create(const orc_graph* g, const orc_params* p);
void orc_destroy(orc_state* st);

/* Set labels (0..B-1) and rebuild M and degrees. */
int orc_set_labels(orc_state* st, const int32_t* labels, int32_t B);
int32_t orc_num_blocks(const orc_state* st);
void orc_get_labels(const orc_state* st, int32_t* out);
/* Degrees d_out/d_in [B]. */
void orc_get_degrees(const orc_state* st, int64_t* d_out, int64_t* d_in);
/* Dense copy of M (B*B int64); only for small B. */
int orc_get_M_dense(const orc_state* st, int64_t* out);

/* data_S as exact fixed point (hi, lo words of the int128) and S = model_S
 * + data_S as double. */
void orc_entropy(const orc_state* st, int64_t* data_hi, uint64_t* data_lo, double* S);

/* One merge proposal of block r, proposal index idx. */
void orc_merge_propose_one(orc_state* st, int32_t r, int32_t idx, uint64_t stream,
                           int32_t* s, int64_t* dS_hi, uint64_t* dS_lo);
/* All blocks, K proposals each; best s and dS per block.  Returns 0, or -1
 * (nothing done) when a pre-drawn stream has fewer than B*K draws left. */
int orc_block_merge(orc_state* st, uint64_t stream, int32_t* best_s,
                    int64_t* dS_hi, uint64_t* dS_lo);
/* The K proposals of blocks [r0, r1) only (best per block kept in the
 * state, no output); times the CPU baseline on a bounded sample. */
void orc_merge_eval(orc_state* st, int32_t r0, int32_t r1, uint64_t stream);
/* Greedy carry out of B_to_merge merges over the last orc_block_merge result;
 * relabels and rebuilds M.  Returns 0, or -1 if the candidates run out. */
int orc_carry_out(orc_state* st, int32_t B_to_merge);

/* One nodal proposal of node n against the current (frozen) state.
 * inf: 0 finite, -1 -inf, +1 +inf, 2 NaN.  accept = (dS < 0). */
void orc_nodal_propose_one(orc_state* st, int32_t n, uint64_t stream, int32_t* s,
                           int64_t* dS_hi, uint64_t* dS_lo, int32_t* inf,
                           int32_t* accept);
/* Nodes [beg,end) against the frozen state, then apply the accepted moves and
 * rebuild M.  Per-node results go to the optional arrays (length end-beg).
 * Returns the sum of accepted dS (fixed point + inf flag) and the move count.
 * Returns 0, or -1 (nothing done) when a pre-drawn stream has fewer than
 * end-beg draws left. */
int orc_nodal_batch(orc_state* st, int32_t beg, int32_t end, uint64_t stream,
                    int32_t* s_out, int64_t* dS_hi, uint64_t* dS_lo, int32_t* inf_out,
                    int32_t* acc_out, int64_t* sum_hi, uint64_t* sum_lo,
                    int32_t* sum_inf, int32_t* n_moves);

/* Same as orc_nodal_batch but only evaluates proposals (no apply); used to
 * time the CPU baseline on a bounded sample. */
void orc_nodal_eval(orc_state* st, int32_t beg, int32_t end, uint64_t stream,
                    int32_t* n_accept);

/* A partition snapshot as the golden-section search sees it: block count,
 * entropy and an id naming which partition it is (B = 0, S = +inf, id = -1
 * for an empty slot, as OldData's constructor, graph_parallel.hpp:116-123). */
typedef struct {
  int64_t B;
  double S;
  int64_t id;
} orc_snap;

/* Golden-section bracket update (graph_parallel.hpp:1322-1365).
 * slots[0..2] = {large, med, small}; cur = the partition just finished.
 * Updates the slots, writes the partition to continue from (*next) and its
 * B_to_merge.  Returns 1 if the optimal B is found (*next is the answer),
 * 0 to continue.  When small is set while large never was (the entropy rose
 * on the second halving) the reference's size_t arithmetic is undefined; the
 * search then stops at med (returns 1). */
int orc_golden_update(orc_snap slots[3], orc_snap cur, double rate, orc_snap* next,
                      int64_t* B_to_merge);

/* Full partition() restatement from the state's current labels (set them with
 * orc_set_labels: SCC or identity init).  Writes final labels; returns B, or
 * a negative error.  stats (optional, ORC_NSTATS doubles): final S, phases,
 * sweeps, batches, nodal proposals, accepted moves, merge proposals, error. */
#define ORC_NSTATS 8
int32_t orc_partition(orc_state* st, int32_t* labels_out, double* stats);

/* One merge phase of the last orc_partition: B after the merge, the MDL at
 * the phase's end, its nodal sweeps, batches and accepted moves. */
typedef struct {
  int64_t B;
  double S;
  int64_t sweeps, batches, moves;
} orc_phase;
/* Copies up to cap phase records of the last orc_partition; returns the count. */
int32_t orc_partition_log(const orc_state* st, int32_t cap, orc_phase* out);

/* The nodal phase of orc_partition (graph_parallel.hpp:381-478) driven by a
 * script instead of a partition state, so that its loop semantics can be
 * checked against hand-derived cases: S_start[i] is the MDL the i-th sweep
 * starts from, and batch j of the phase returns the accepted-dS sum
 * (sum_hi/lo[j], fixed point; sum_inf[j] 0 or -1) and moves[j] (nullable).
 * out[6] = sweeps, batches, proposals, moves, converged, S after the last
 * batch; itr_dS (nullable, [max_nodal_itr]) = itr_delta_entropy.  Returns -1
 * if the script ran out, else 0. */
int orc_nodal_phase_scripted(const orc_params* p, int32_t N, int32_t batch, int32_t isf,
                             const double* S_start, int64_t n_S, const int64_t* sum_hi,
                             const uint64_t* sum_lo, const int32_t* sum_inf,
                             const int32_t* moves, int64_t n_batch, double* out,
                             double* itr_dS);

/* Pre-drawn uniform stream in place of Philox (u: n/4 draws of four
 * uniforms in [0,1), slot order as above; a uniform u stands for the word
 * floor(u 2^64)).  Draws are taken from a cursor, reset to 0 here:
 *   orc_block_merge: block r, proposal j takes draw pos + r*K + j; the round
 *     takes B*K draws;
 *   orc_nodal_batch(beg, end): node n takes draw pos + (n - beg); the batch
 *     takes end - beg draws (so the batches of a sweep, in order, give node n
 *     draw pos_sweep + n);
 *   orc_merge_propose_one / orc_merge_eval read the next round's draws and
 *   orc_nodal_propose_one (node n: pos + n) / orc_nodal_eval the next
 *   batch's, without taking them.
 * Stream arguments are ignored meanwhile; orc_partition returns -4 if the
 * stream runs out.  u = NULL, n = 0 returns to Philox.  Returns 0, or -1 for
 * a bad buffer (n not a multiple of 4, a value outside [0,1)). */
int orc_set_uniforms(orc_state* st, const double* u, int64_t n);
/* Draws consumed since orc_set_uniforms. */
int64_t orc_uniforms_used(const orc_state* st);

/* Textbook double-precision data_S = -sum m*log(m/(d_out*d_in)) (no fixed
 * point) for cross-checks. */
double orc_data_S_double(const orc_state* st);
/* model_S of graph_parallel.hpp:1314-1316 in double. */
double orc_model_S(int64_t N, int64_t E, int64_t B);

#ifdef __cplusplus
}
#endif
#endif
