/*
 * usap.h -- C-ABI of the MI355X stochastic block partitioner engine.
 *
 * Plain C types only (pointers and sizes); no C++ or torch types cross this
 * boundary.  All functions return USAP_OK (0) or a negative USAP_E* code and
 * never throw or exit.  One engine handle owns one HIP stream and all device
 * memory of one graph; a handle is used from one host thread at a time, and
 * several handles may run concurrently.
 *
 * The engine is the drop-in for the hot path of sgp::Graph_P<W>::partition()
 * in chih-chun-chang/uSAP include/graph_parallel.hpp (the reference has no
 * FFI; its path sits behind private member functions called from Taskflow
 * tasks).  Each entry point names the reference function it replaces.
 *
 * Arithmetic shared with the CPU oracle:
 *   lambda(x) = llrint(log(x) * 2^48), T(m,x,y) = m*(lambda(m)-lambda(x)-lambda(y))
 *   summed exactly in int128 (scale 2^-48); a dS is reported as a double and,
 *   through the *_fx entry points, as its exact int128 (hi, lo) words.
 * Random stream: Philox4x64-10, key {seed, 0}, counter {item, proposal,
 *   stream, tag}; slots 0 neighbour, 1 accept-uniform, 2 uniform block,
 *   3 multinomial; an index in [0, W) is (x * W) >> 64.  usap_set_uniforms
 *   replaces it with a caller's pre-drawn stream.
 */
#ifndef USAP_H
#define USAP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define USAP_OK 0
#define USAP_EINVAL (-1)       /* bad argument */
#define USAP_ENOMEM (-2)       /* host or device allocation failed */
#define USAP_EHIP (-3)         /* HIP runtime error (no device, launch failure) */
#define USAP_ESTATE (-4)       /* call not valid in this state (e.g. golden-section bracket invalid) */
#define USAP_EINFEASIBLE (-5)  /* carry-out ran out of candidate merges */
#define USAP_EIO (-6)          /* file could not be read or parsed */

#define USAP_INIT_SCC 0        /* graph_parallel.hpp:514-556 (reference default) */
#define USAP_INIT_IDENTITY 1   /* B starts at N */
#define USAP_DS_REF 0          /* nodal dS of graph_parallel.hpp:1066-1215 */
#define USAP_DS_EXACT 1        /* nodal dS with the corner in-edge updates of graph.cu:1414-1426 */

/* A directed graph as CSR (out-edges) + CSC (in-edges), both in the
 * reference's neighbour order (file order, graph_parallel.hpp:290-294).
 * Ids are 0-based. */
typedef struct usap_graph {
  int32_t N;
  int64_t E;
  const int64_t* out_ptr; /* [N+1] */
  const int32_t* out_idx; /* [E] */
  const int64_t* in_ptr;  /* [N+1] */
  const int32_t* in_idx;  /* [E] */
  const int32_t* out_w;   /* [E], NULL = unweighted (all 1) */
  const int32_t* in_w;    /* [E], NULL = unweighted */
} usap_graph;

/* Parameters of Graph_P (graph_parallel.hpp:41-52) plus engine options. */
typedef struct usap_params {
  int32_t K;                       /* num_agg_proposals_per_block = 10 */
  double reduction_rate;           /* num_block_reduction_rate = 0.5 */
  int32_t max_nodal_itr;           /* max_num_nodal_itr = 100 */
  int32_t num_batch_nodal_update;  /* 4: nodal batch = N / 4 */
  int32_t nodal_batch;             /* > 0: fixed nodal batch size (overrides the above) */
  double thr1;                     /* delta_entropy_threshold1 = 5e-4 */
  double thr2;                     /* delta_entropy_threshold2 = 1e-4 */
  int32_t avg_window;              /* delta_entropy_moving_avg_window = 3 */
  int32_t dfs_depth;               /* dfs_depth = 20 */
  int32_t init;                    /* USAP_INIT_SCC | USAP_INIT_IDENTITY */
  int32_t ds_mode;                 /* USAP_DS_REF | USAP_DS_EXACT */
  uint64_t seed;                   /* random-stream key */
  int32_t device;                  /* HIP device ordinal */
  int32_t verbose;                 /* print per-phase progress to stderr */
  int32_t slab_slack;              /* hashed block row capacity = pow2 >= slack*(out+in edges)+16; a row grows past 1/slack load (2; 1..16) */
  int32_t graphs;                  /* 1: replay each nodal sweep as a captured HIP graph (0 =
                                      default).  The capture is thread-local
                                      (hipStreamCaptureModeThreadLocal).  Use it with one
                                      engine at a time per process: usap_partition_batch and
                                      usap_serve refuse it (USAP_EINVAL) with more than one
                                      worker, since a process replaying six engines' sweeps
                                      on six host threads once aborted (cause not found) */
} usap_params;

/* Partition statistics (usap_partition).  Times are host wall-clock ms. */
typedef struct usap_stats {
  double S;                 /* final MDL */
  int32_t B;                /* final block count */
  int32_t phases;           /* merge phases run */
  int64_t sweeps;           /* nodal sweeps */
  int64_t batches;          /* nodal batches */
  int64_t nodal_proposals;  /* nodal proposals evaluated (node-moves) */
  int64_t accepted_moves;   /* accepted nodal moves */
  int64_t merge_proposals;  /* block-merge proposals evaluated */
  double ms_total;          /* partition() window, graph upload excluded */
  double ms_merge;          /* merge proposals */
  double ms_carry;          /* carry-out + relabel + M rebuild */
  double ms_nodal;          /* nodal batches (propose + apply) */
  double ms_entropy;        /* MDL evaluations */
  double ms_golden;         /* snapshot / restore */
  int32_t rebuilds;         /* full M rebuilds */
  int32_t error;            /* 0 or USAP_E* */
  int64_t grows;            /* M rows moved to a larger slab between batches */
} usap_stats;

typedef struct usap_engine usap_engine;

/* Defaults of graph_parallel.hpp:41-52 (seed 0, device 0, SCC init, ref dS). */
void usap_default_params(usap_params* p);

/* Create an engine: uploads the graph, builds the lambda table, allocates
 * device state for up to N blocks.  ≙ Graph_P ctor (graph_parallel.hpp:61-78)
 * minus the file load (see usap_load_tsv). */
int usap_create(const usap_graph* g, const usap_params* p, usap_engine** out);
int usap_destroy(usap_engine* e);
const char* usap_last_error(const usap_engine* e);

/* Replace the engine's edges (same N), keeping the labels and B; M and the
 * block degrees are rebuilt for the new edges.  No counterpart in the
 * reference (one graph per Graph_P object, graph_parallel.hpp:61-78); for
 * callers that receive a graph as streamed edge increments. */
int usap_set_graph(usap_engine* e, const usap_graph* g);

/* Appends n edges src[i] -> dst[i] (weights w, NULL = 1; weights only for
 * a weighted graph) after the graph's current edges, keeping the labels and
 * B.  Only the increment crosses to the device: the CSR/CSC lists grow on
 * the device, each node's new neighbours after its old ones in delivery
 * order (the reference's neighbour order, graph_parallel.hpp:290-294), and
 * M and the block degrees are rebuilt.  Equal to usap_set_graph with the
 * CSR/CSC of the concatenated edge list (usap_build_csr).  No counterpart
 * in the reference; for callers that receive a graph as streamed edge
 * increments (its ingest, load_graph_from_tsv, graph_parallel.hpp:251-317,
 * reads one whole file). */
int usap_append_edges(usap_engine* e, int64_t n, const int32_t* src, const int32_t* dst,
                      const int32_t* w);

/* usap_create from an edge list in delivery order (0-based; w NULL =
 * unweighted): the CSR/CSC are built on the device.  ≙ Graph_P ctor +
 * load_graph_from_tsv's list building (graph_parallel.hpp:61-78, :251-317)
 * on an edge list already in memory. */
int usap_create_from_edges(int32_t N, int64_t E, const int32_t* src, const int32_t* dst,
                           const int32_t* w, const usap_params* p, usap_engine** out);

/* Set labels (all in [0,B)) and rebuild M and the block degrees.
 * ≙ labels assignment + _initialize_edge_counts (graph_parallel.hpp:703-759). */
int usap_set_labels(usap_engine* e, const int32_t* labels, int32_t B);
int usap_get_labels(usap_engine* e, int32_t* out);
int usap_num_blocks(usap_engine* e, int32_t* B);
/* Block degrees d_out/d_in [B]. ≙ Partition::d (graph_parallel.hpp:85-99). */
int usap_get_degrees(usap_engine* e, int32_t* d_out, int32_t* d_in);
/* M as COO triplets sorted by (row, col), zero entries omitted; *nnz gets the
 * count (call with cap 0 to query).  ≙ Partition::M / Mrow (graph_parallel.hpp:101-110). */
int usap_get_M_coo(usap_engine* e, int64_t cap, int32_t* rows, int32_t* cols, int32_t* vals,
                   int64_t* nnz);

/* MDL S = model_S + data_S.  ≙ _compute_overall_entropy (graph_parallel.hpp:1281-1319).
 * data_fx_hi/lo (nullable): exact data_S as int128 words. */
int usap_entropy(usap_engine* e, double* S, int64_t* data_fx_hi, uint64_t* data_fx_lo);

/* usap_entropy without any allocation or host wait.  No counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops).  Launches on the engine's stream and writes the MDL
 * of the current state to device memory the caller owns: S_dev (one double,
 * nullable) and fx_dev (the exact data term as int128 words hi, lo;
 * nullable); the values equal usap_entropy's once the stream has reached
 * them.  Partial sums live in a 64 KiB static device array (4096 partials,
 * the most the grid-strided scan uses), one per GPU, shared by every call on
 * that GPU and holding the last call's contents: calls that overlap on one
 * GPU race.  USAP_EINVAL if both outputs are NULL. */
int usap_entropy_noalloc(usap_engine* e, double* S_dev, int64_t* fx_dev);

/* K merge proposals per block and the best per block.
 * ≙ block_merge task (graph_parallel.hpp:349-368) over
 *   _propose_new_partition_block (graph_parallel.hpp:761-988).
 * best_s/best_dS (nullable, [B]) receive the result. */
int usap_merge_propose(usap_engine* e, int32_t K, uint64_t stream, int32_t* best_s,
                       double* best_dS);
/* Exact dS words of the last usap_merge_propose ([B] each). */
int usap_merge_result_fx(usap_engine* e, int32_t* best_s, int64_t* dS_hi, uint64_t* dS_lo);

/* Carry out B_to_merge merges of the last proposal round, relabel, compact,
 * rebuild M.  ≙ _carry_out_best_merges (graph_parallel.hpp:1225-1279) +
 * _initialize_edge_counts.  USAP_EINFEASIBLE if the candidates run out
 * (undefined behaviour in the reference); the state is then unchanged. */
int usap_merge_apply(usap_engine* e, int32_t B_to_merge, int32_t* B_new);

/* One nodal batch over nodes [begin,end): proposals against the frozen M,
 * greedy accept (dS < 0), then the accepted moves applied to labels, M and
 * degrees.  ≙ nodal_move + update tasks (graph_parallel.hpp:406-458) over
 * _propose_new_partition_nodal (graph_parallel.hpp:990-1222).
 * dS_sum = sum of accepted dS (may be -inf in ref mode). */
int usap_nodal_batch(usap_engine* e, int32_t begin, int32_t end, uint64_t stream,
                     double* dS_sum, int32_t* n_moves);
/* One nodal sweep: batches [0,b), [b,2b), ... of size `batch`, the i-th with
 * random stream stream0 + i, each followed on the device by the reference's
 * update step (graph_parallel.hpp:433-458): the sweep's cumulative accepted
 * dS (dS_sum), S += dS_sum, and, when `check` is set (itr >= window-1), the
 * convergence test -mean < thr * S, which ends the sweep.  Equal to calling
 * usap_nodal_batch batch by batch with that test on the host; one host
 * synchronisation per sweep. */
typedef struct usap_sweep {
  int32_t batches;     /* batches executed */
  int32_t converged;   /* the convergence test fired after the last one */
  int64_t proposals;   /* nodal proposals evaluated */
  int64_t moves;       /* accepted moves */
  double S;            /* S after the last executed batch */
  double dS_sum;       /* cumulative accepted dS of the sweep (itr_delta_entropy[itr]) */
  int64_t dS_hi;       /* the same, exact int128 words */
  uint64_t dS_lo;
  int32_t dS_inf;      /* -1 when an accepted dS was -inf */
  int32_t rebuilds;    /* full M rebuilds after slab overflows */
} usap_sweep;
int usap_nodal_sweep(usap_engine* e, int32_t batch, uint64_t stream0, double S, double mean,
                     double thr, int32_t check, usap_sweep* out);

/* Deterministic harness mode: the engine draws from a caller's pre-drawn
 * uniform stream instead of Philox.  Replaces the per-worker generators
 * (_pt_generator, graph_parallel.hpp:67-77, :180) at their draw sites
 * (merge :805-862, nodal :1022-1046), so that the reference, fed the same
 * buffer, makes the same proposals.  u[n]: n/4 draws of four uniforms in
 * [0, 1), slot order as the Philox words (neighbour, accept-uniform,
 * uniform block, multinomial); a uniform u stands for the word floor(u 2^64),
 * so an index in [0, W) is floor(u W) and the uniform branch tests
 * u' <= B / (d + B) with u' = u rounded down to a multiple of 2^-53.
 * Draws are consumed in call order from a cursor (reset to 0 here):
 *   usap_merge_propose(K): block r, proposal j takes draw pos + r*K + j;
 *     the round takes B*K draws;
 *   usap_nodal_batch(begin, end): node n takes draw pos + (n - begin); the
 *     batch takes end - begin draws;
 *   usap_nodal_sweep: node n takes draw pos + n; the sweep takes one draw per
 *     proposal evaluated (N if it runs to the end) and needs N available;
 * so a sweep equals its batches called one by one, and usap_partition /
 * usap_resume consume the stream in the reference's own order.  The stream
 * arguments of those calls are ignored while the mode is on.  USAP_ESTATE
 * (nothing run, cursor unchanged) when a call needs more draws than are left.
 * u = NULL, n = 0 returns to Philox.  Copies u to the device. */
int usap_set_uniforms(usap_engine* e, const double* u, size_t n);
/* Draws consumed from the pre-drawn stream since usap_set_uniforms. */
int usap_uniforms_used(usap_engine* e, int64_t* draws);

/* Exact sum of the last batch: int128 words and inf flag (0 or -1). */
int usap_nodal_sum_fx(usap_engine* e, int64_t* hi, uint64_t* lo, int32_t* inf);
/* Per-node results of the last batch ([end-begin] each, nullable):
 * proposed block, dS words, inf flag (0 finite, -1 -inf), accepted. */
int usap_nodal_detail(usap_engine* e, int32_t* s, int64_t* dS_hi, uint64_t* dS_lo, int32_t* inf,
                      int32_t* accept);

/* Snapshot slots (0..2) hold labels, B and S.  ≙ OldData copies
 * (graph_parallel.hpp:111-124, :1327-1362). */
int usap_snapshot(usap_engine* e, int32_t slot, double S);
int usap_restore(usap_engine* e, int32_t slot);

/* Full partition(): init (params.init), then merge phases, nodal sweeps and
 * the golden-section search over B, run by the host C++ driver over the
 * entry points above.  ≙ Graph_P::partition (graph_parallel.hpp:322-512) +
 * _prepare_for_partition_next (graph_parallel.hpp:1322-1365).
 * labels_out (nullable, [N]); stats (nullable). */
int usap_partition(usap_engine* e, int32_t* labels_out, usap_stats* stats);

/* The nodal phase (graph_parallel.hpp:381-478) at the current labels and B:
 * sweeps of batches from nodal stream stream0 until the convergence test
 * with delta_entropy_threshold2 fires (or max_nodal_itr sweeps), then the
 * MDL.  No merge phase and no golden section.  No counterpart in the
 * reference; for callers that resume the chain on a grown graph
 * (usap_set_graph) from the previous partition.  labels_out / stats
 * nullable; stats->S and ->B are the final MDL and B. */
int usap_resume(usap_engine* e, uint64_t stream0, int32_t* labels_out, usap_stats* stats);

/* Per-phase record of the last usap_partition (one per merge phase). */
typedef struct usap_phase {
  int32_t B;           /* blocks after the merge */
  int32_t sweeps;      /* nodal sweeps run */
  int64_t proposals;   /* nodal proposals evaluated */
  int64_t moves;       /* accepted nodal moves */
  double S;            /* MDL at the end of the phase */
  double ms_merge;     /* merge proposals + carry-out */
  double ms_nodal;     /* nodal sweeps */
  double k_nodal_ms;   /* kernel times in this phase (profiling on, else 0) */
  double k_apply_ms;
  double k_merge_ms;
  double k_build_ms;
  double k_grow_ms;
  int32_t rebuilds;    /* full M rebuilds in this phase */
  int64_t merge_proposals; /* K x blocks before the merge */
} usap_phase;

/* Copies up to cap phase records of the last usap_partition; *n = count. */
int usap_partition_log(usap_engine* e, int32_t cap, usap_phase* out, int32_t* n);

/* ---- many graphs at once ---------------------------------------------------
 * No counterpart in the reference, which partitions one graph per process
 * (main.cpp:48-63) with partition() (graph_parallel.hpp:322-512) as the
 * per-graph state machine.  For callers with many independent graphs. */

typedef struct usap_batch_stats {
  int32_t workers;         /* graphs in flight at once (one engine + HIP stream each) */
  int32_t failed;          /* graphs whose partition failed (stats[i].error != 0) */
  double ms_total;         /* wall time of the whole batch */
  double graphs_per_s;
  int64_t nodal_proposals; /* over all graphs (when stats is given) */
} usap_batch_stats;

/* Partitions graphs[0..n) (params[i] for graph i, as usap_create +
 * usap_partition) with `workers` graphs in flight at once, each on its own
 * engine and HIP stream, largest graphs first.  Graph i writes its N_i labels
 * to labels + label_off[i] and its statistics to stats[i] (nullable).  Each
 * graph's result is the one it gets alone.  workers = 0: chosen by
 * usap_batch_autotune.  USAP_EINVAL if a graph asks for captured-sweep
 * replay (params[i].graphs) and workers != 1.  Returns USAP_ESTATE if any
 * graph failed (see stats[i].error and out->failed). */
int usap_partition_batch(int32_t n, const usap_graph* graphs, const usap_params* params,
                         int32_t* labels, const int64_t* label_off, usap_stats* stats,
                         int32_t workers, usap_batch_stats* out);

/* ---- serving graphs that arrive as streamed edge increments ---------------
 * No counterpart in the reference (streaming is future work there,
 * README.md:47-49).  For callers that serve partition requests over time. */

typedef struct usap_serve_graph {
  int32_t N;
  int64_t E;
  const int32_t* src;   /* [E] 0-based edge sources in delivery order */
  const int32_t* dst;   /* [E] */
  const int32_t* w;     /* [E] weights, NULL = unweighted */
  double arrival_s;     /* first increment's arrival, seconds after the start */
  usap_params params;
} usap_serve_graph;

typedef struct usap_serve_opts {
  int32_t stages;       /* increments per graph: stage k holds the first ceil(E*k/stages) edges */
  double interval_s;    /* time between a graph's increments */
  int32_t workers;      /* stages in flight at once (one host thread each) */
  int64_t hbm_budget;   /* device bytes for resident graphs by usap_footprint (0 = no limit) */
  int32_t final_partition; /* 1: the last stage runs a full partition of the whole graph
                              instead of resuming (the resumed chain keeps stage 1's B) */
  double repartition_growth; /* > 1: a stage whose edges reach this multiple of the
                                edges at the graph's last full partition runs a full
                                partition (merges and the search over B) instead of
                                resuming; 0 = never */
} usap_serve_opts;

typedef struct usap_serve_stage {  /* one per (graph, stage), graph-major */
  int32_t graph, stage;
  double t_ready;       /* increment arrived (s since the start) */
  double t_start, t_done;
  double t_admit;       /* stage 1 only: graph admitted */
  double t_release;     /* stage 1 only: graph left the device */
  double S;             /* MDL after this stage */
  int32_t B;
  int32_t error;
  int64_t E;            /* edges in this stage */
  int64_t sweeps;
  int64_t nodal_proposals;
  int32_t full;         /* 1: this stage ran a full partition */
} usap_serve_stage;

typedef struct usap_serve_stats {
  double wall_s;
  double graphs_per_s;  /* graphs / (last completion - first arrival) */
  double p50_ms, p99_ms;                /* stage latency t_done - t_ready, all stages */
  double p50_first_ms, p99_first_ms;    /* stage 1 (full partition) */
  double p50_resume_ms, p99_resume_ms;  /* stages 2.. (resumed chain) */
  int32_t max_resident; /* graphs resident at once */
  int32_t failed;
  int64_t max_bytes;    /* their summed footprint at the peak */
} usap_serve_stats;

/* Estimated device bytes of an engine for a graph of N nodes and E edges
 * (graph arrays, state, bucket sums and M slabs at B = N), used for
 * admission. */
int64_t usap_footprint(int32_t N, int64_t E);

/* Serves n graphs: graph g arrives at graphs[g].arrival_s and its stage k
 * (1..stages) becomes ready interval_s*(k-1) later.  Stage 1 =
 * usap_create_from_edges + usap_partition; stage k > 1 = usap_append_edges
 * (the increment only) + usap_resume (nodal streams from k << 40), so the
 * engine stays resident between stages; a stage runs usap_partition instead
 * when the graph has grown by repartition_growth since its last full
 * partition, and so does the last stage with final_partition.
 * Graphs are admitted in arrival order while their footprints fit
 * hbm_budget; up to `workers` ready stages of different graphs run at once,
 * oldest first.  Each graph's final labels go to labels + label_off[g];
 * stages[g*stages + k-1] records stage k.  Every result equals the one the
 * graph gets served alone.  USAP_EINVAL if a graph asks for captured-sweep
 * replay (params.graphs) with more than one worker.  USAP_ESTATE if a graph
 * failed (its stage record holds the error). */
int usap_serve(int32_t n, const usap_serve_graph* graphs, const usap_serve_opts* opt,
               int32_t* labels, const int64_t* label_off, usap_serve_stage* stages,
               usap_serve_stats* out);

/* Picks the number of graphs in flight: partitions a sample of `sample`
 * graphs spread over the size range (0 = 24) with 1, 2, 4, 8, ... workers (at
 * most max_workers, 0 = 16) and takes the count with the best graphs/s.  tried /
 * rates (nullable, >= 16 entries) receive each trial's workers and graphs/s.
 * USAP_EINVAL if a graph asks for captured-sweep replay and max_workers != 1. */
int usap_batch_autotune(int32_t n, const usap_graph* graphs, const usap_params* params,
                        int32_t max_workers, int32_t sample, int32_t* workers, int32_t* tried,
                        double* rates);

/* Kernel profile: time on the engine's stream (nodal, apply and grow: from
 * the launch's first workgroup start to its last workgroup end, stamped in
 * the kernel; the others: events around their launches), launches, the
 * bytes the kernel moved by its own algorithm (reads + writes it needs, each
 * counted once; 0 where not tallied), and, for the nodal and merge kernels,
 * the compulsory bytes of the reference's algorithm for the same work, from
 * the run's own counts: per nodal batch Q
 *   A(Q) = sum_n (16 + 4 + 8 k_n + 12 [+ 4 k_n weighted])
 *        + sum_{b in D(Q)} (16 + 8 nnzR(b) + 8 nnzC(b) + 12)
 *        + sum_{n accepted} (4 + 16 k_n),
 * D(Q) the distinct blocks among the batch's r_n, s_n, u_n (old block,
 * proposed block, drawn neighbour's block; graph_parallel.hpp:990-1222 reads
 * their rows and columns of M whole); per merge round
 *   A = 2 (4 (B+1) + 8 nnz(M)) + 12 B + 12 B
 * (graph_parallel.hpp:761-988, rows and columns of M read once).  Counting
 * them runs extra kernels, so profiled times come from a separate run. */
#define USAP_K_NODAL 0   /* nodal proposals */
#define USAP_K_APPLY 1   /* accepted-move application */
#define USAP_K_MERGE 2   /* merge proposals */
#define USAP_K_ENTROPY 3 /* MDL reduction */
#define USAP_K_BUILD 4   /* M rebuild (all its kernels) */
#define USAP_K_GROW 5    /* row growth / overflow replay between nodal batches */
#define USAP_NKERNELS 6
typedef struct usap_kernel_prof {
  int64_t launches;
  double ms;
  double bytes;        /* this kernel's own algorithmic bytes */
  double model_bytes;  /* the reference algorithm's compulsory bytes (nodal, merge; else 0) */
} usap_kernel_prof;
/* enable = 1: reset and start timing; 0: stop; -1: leave as is.  out
 * (nullable, [USAP_NKERNELS]) receives the totals since the last reset. */
int usap_profile(usap_engine* e, int32_t enable, usap_kernel_prof* out);

/* Golden-section bracket update on (B, S, id) triples, exposed for tests.
 * ≙ _prepare_for_partition_next (graph_parallel.hpp:1322-1365).
 * slots_* [3] = large, med, small (B = 0, S = +inf, id = -1 when empty).
 * Returns 1 if the optimal B is found (also when the entropy rose on the
 * second halving, where the reference is undefined: the search stops at med),
 * 0 to continue. */
int usap_golden_update(int64_t* slot_B, double* slot_S, int64_t* slot_id, int64_t cur_B,
                       double cur_S, int64_t cur_id, double rate, int64_t* next_B,
                       double* next_S, int64_t* next_id, int64_t* B_to_merge);

/* ---- host utilities (no device needed) ---------------------------------- */

/* SCC initial partition.  ≙ _scc + _dfs_itr (graph_parallel.hpp:514-603).
 * Reads the out-lists; the in-lists (optional, NULL allowed) are used as the
 * transpose when each lists its sources in ascending order with the
 * out-lists' in-degrees, else the transpose is built from the out-lists. */
int usap_scc_init(const usap_graph* g, int32_t dfs_depth, int32_t* labels, int32_t* B);

/* Model entropy term.  ≙ graph_parallel.hpp:1314-1316 in double. */
double usap_model_S(int64_t N, int64_t E, int64_t B);

/* lambda table [maxv+1] (host copy of what the engine uploads). */
int usap_lambda_table(int64_t maxv, int64_t* out);

/* Philox4x64-10 on the host: out[4] = philox(ctr[4], key[2]). */
void usap_philox4x64(const uint64_t* ctr, const uint64_t* key, uint64_t* out);

/* Build CSR/CSC from an edge list in file order (stable by source / target).
 * src/dst 0-based; w nullable.  Output arrays sized by the caller:
 * out_ptr/in_ptr [N+1], out_idx/in_idx [E], out_w/in_w [E] (nullable). */
int usap_build_csr(int32_t N, int64_t E, const int32_t* src, const int32_t* dst,
                   const int32_t* w, int64_t* out_ptr, int32_t* out_idx, int32_t* out_w,
                   int64_t* in_ptr, int32_t* in_idx, int32_t* in_w);

/* Read a GraphChallenge TSV ("from\tto\tweight", 1-based).  ≙ load_graph_from_tsv
 * (graph_parallel.hpp:251-317), except N = max id over both columns.  Call
 * with NULL arrays to query N and E, then again with arrays of that size. */
int usap_load_tsv(const char* path, int32_t* N, int64_t* E, int32_t* src, int32_t* dst,
                  int32_t* w);

/* Seeded directed degree-corrected SBM: K planted blocks with sizes
 * min_block + Dirichlet(alpha) share of the rest, out-degree power law
 * d^-gamma on [dmin, dmax], a fraction p_in of edges inside the block and the
 * rest to a uniformly chosen other block, no self-loops or duplicates.
 * Edges come out grouped by source ascending.  Call with NULL src/dst/truth
 * to query E (deterministic), then again with arrays. */
int usap_generate_dcsbm(int32_t N, int32_t K, double alpha, int32_t min_block, int32_t dmin,
                        int32_t dmax, double gamma, double p_in, uint64_t seed, int64_t* E,
                        int32_t* src, int32_t* dst, int32_t* truth);

#ifdef __cplusplus
}
#endif
#endif
