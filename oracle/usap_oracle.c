/*
 * usap_oracle.c -- CPU restatement of uSAP's partition hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the CPU baseline.  The
 * product never links or calls this file.  See usap_oracle.h for the list of
 * reference functions restated here and for the arithmetic and random-stream
 * definitions.
 *
 * Complexity follows the reference on purpose: dense M when B < block_size,
 * else per-block edge lists with duplicates (graph_parallel.hpp:703-759);
 * O(B) dense scratch rows per proposal (graph_parallel.hpp:790, :1072);
 * M rebuilt from scratch after every nodal batch (graph_parallel.hpp:443);
 * O(B_to_merge * (B + N)) carry-out (graph_parallel.hpp:1240-1257).
 */
#define _GNU_SOURCE
#include "usap_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

typedef __int128 i128;
typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------ */
/* Random stream                                                             */
/* ------------------------------------------------------------------------ */

/* Philox4x64-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11: multipliers and Weyl increments of the paper. */
void orc_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
  uint64_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint64_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    if (round > 0) {
      k0 += 0x9E3779B97F4A7C15ull;
      k1 += 0xBB67AE8584CAA73Bull;
    }
    u128 p0 = (u128)0xD2E7470EE14C6C93ull * c0;
    u128 p1 = (u128)0xCA5A826395121157ull * c2;
    uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0;
    uint64_t hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
    uint64_t n0 = hi1 ^ c1 ^ k0;
    uint64_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

static void draw4(uint64_t seed, uint64_t item, uint64_t prop, uint64_t stream,
                  uint64_t tag, uint64_t x[4]) {
  uint64_t ctr[4] = {item, prop, stream, tag};
  uint64_t key[2] = {seed, 0};
  orc_philox4x64(ctr, key, x);
}

/* index in [0, W): floor(x * W / 2^64) */
static inline int64_t idx_of(uint64_t x, int64_t W) {
  return (int64_t)(((u128)x * (u128)(uint64_t)W) >> 64);
}

/* accept-uniform test of graph_parallel.hpp:816/:1034,
 * x <= B / (d_a[u] + B) with x = (x1 >> 11) * 2^-53, evaluated exactly. */
static inline int uniform_branch(uint64_t x1, int64_t da_u, int64_t B) {
  u128 lhs = (u128)(x1 >> 11) * (u128)(uint64_t)(da_u + B);
  u128 rhs = (u128)(uint64_t)B << 53;
  return lhs <= rhs;
}

/* uniform block different from r (graph_parallel.hpp:817-821) */
static inline int32_t uniform_not_r(uint64_t x2, int32_t r, int32_t B) {
  int32_t s = (int32_t)idx_of(x2, B);
  if (s == r) s++;
  if (s == B) s = 0;
  return s;
}

int64_t orc_pick(const int64_t* w, int64_t n, int64_t t) {
  int64_t acc = 0;
  for (int64_t i = 0; i < n; ++i) {
    acc += w[i];
    if (acc > t) return i;
  }
  return n - 1; /* unreachable when t < sum w */
}

/* ------------------------------------------------------------------------ */
/* Fixed-point entropy terms                                                 */
/* ------------------------------------------------------------------------ */

void orc_lambda_table(int64_t maxv, int64_t* out) {
  out[0] = 0;
  for (int64_t x = 1; x <= maxv; ++x) out[x] = llrint(log((double)x) * 281474976710656.0);
}

typedef struct {
  i128 fx;
  int inf; /* 0 finite, -1 -inf, +1 +inf, 2 NaN */
} dsacc;

static inline void acc_inf(dsacc* a, int sgn) {
  if (a->inf == 2) return;
  if (a->inf == 0)
    a->inf = sgn;
  else if (a->inf != sgn)
    a->inf = 2;
}

struct tscratch;

struct orc_state {
  orc_graph g;
  orc_params p;
  int64_t* lam;
  int64_t lam_n;
  int32_t B;
  int32_t* labels;
  /* dense M (B < block_size) */
  int64_t* M;
  /* edge lists with duplicates (B >= block_size), graph_parallel.hpp:731-745 */
  int64_t* row_ptr;
  int32_t* row_idx;
  int64_t* row_w;
  int64_t* col_ptr;
  int32_t* col_idx;
  int64_t* col_w;
  int64_t* d_out;
  int64_t* d_in;
  int64_t* d_a;
  /* last block-merge result */
  int32_t* best_s;
  i128* best_dS;
  int32_t cap_B; /* allocation size for per-block arrays */
  /* per-phase record of the last orc_partition */
  orc_phase* plog;
  int32_t plog_n, plog_cap;
  /* pre-drawn uniform stream (orc_set_uniforms): 4 words per draw, or NULL */
  uint64_t* uni;
  int64_t uni_n, uni_pos; /* draws held, draws consumed */
  /* per-worker scratch kept between calls (scratch_get) */
  struct tscratch* ts;
  int ts_threads;
  int32_t ts_B;
};

/* Draw d of the pre-drawn stream (its four words, Philox slot order). */
static void pre_draw(const orc_state* st, uint64_t d, uint64_t x[4]) {
  if (d >= (uint64_t)st->uni_n) abort(); /* callers reserve their draws first */
  for (int j = 0; j < 4; ++j) x[j] = st->uni[4 * d + (uint64_t)j];
}

/* Draws left in the pre-drawn stream. */
static int64_t uni_left(const orc_state* st) { return st->uni_n - st->uni_pos; }

/* sgn * T(m, x, y) */
static inline void acc_T(const orc_state* st, dsacc* a, int sgn, int64_t m, int64_t x,
                         int64_t y) {
  if (m == 0) return;
  if (m < 0 || x < 0 || y < 0) {
    a->inf = 2;
    return;
  }
  if (x == 0 || y == 0) {
    acc_inf(a, sgn);
    return;
  }
  if (m >= st->lam_n || x >= st->lam_n || y >= st->lam_n) abort();
  i128 t = (i128)m * (i128)(st->lam[m] - st->lam[x] - st->lam[y]);
  if (sgn > 0)
    a->fx += t;
  else
    a->fx -= t;
}

static double fx_to_double(i128 v) {
  int neg = v < 0;
  u128 a = neg ? (u128)0 - (u128)v : (u128)v;
  uint64_t hi = (uint64_t)(a >> 64), lo = (uint64_t)a;
  double d = ldexp((double)hi, 64 - ORC_LAMBDA_FRAC) + ldexp((double)lo, -ORC_LAMBDA_FRAC);
  return neg ? -d : d;
}

static double acc_to_double(const dsacc* a) {
  if (a->inf == 2) return NAN;
  if (a->inf == 1) return INFINITY;
  if (a->inf == -1) return -INFINITY;
  return fx_to_double(a->fx);
}

static inline void split128(i128 v, int64_t* hi, uint64_t* lo) {
  *hi = (int64_t)(v >> 64);
  *lo = (uint64_t)v;
}

/* ------------------------------------------------------------------------ */
/* Parallel for (stands in for tf::Executor::for_each_index)                 */
/* ------------------------------------------------------------------------ */

typedef void (*item_fn)(void* ctx, int tid, int64_t item);

typedef struct {
  item_fn fn;
  void* ctx;
  int64_t n;
  atomic_llong next;
  int tid;
} pf_job;

typedef struct {
  pf_job* job;
  int tid;
} pf_arg;

static void* pf_worker(void* vp) {
  pf_arg* a = (pf_arg*)vp;
  pf_job* j = a->job;
  for (;;) {
    int64_t b = atomic_fetch_add(&j->next, 8);
    if (b >= j->n) break;
    int64_t e = b + 8 < j->n ? b + 8 : j->n;
    for (int64_t i = b; i < e; ++i) j->fn(j->ctx, a->tid, i);
  }
  return NULL;
}

static void par_for(int threads, int64_t n, item_fn fn, void* ctx) {
  if (threads <= 1 || n < 64) {
    for (int64_t i = 0; i < n; ++i) fn(ctx, 0, i);
    return;
  }
  pf_job job;
  job.fn = fn;
  job.ctx = ctx;
  job.n = n;
  atomic_init(&job.next, 0);
  pthread_t th[256];
  pf_arg args[256];
  if (threads > 256) threads = 256;
  for (int t = 0; t < threads; ++t) {
    args[t].job = &job;
    args[t].tid = t;
    if (t > 0) pthread_create(&th[t], NULL, pf_worker, &args[t]);
  }
  pf_worker(&args[0]);
  for (int t = 1; t < threads; ++t) pthread_join(th[t], NULL);
}

/* per-worker dense scratch: NewM (graph_parallel.hpp:126-140) + prob */
typedef struct tscratch {
  int64_t *Rr, *Rs, *Cr, *Cs, *prob;
} tscratch;

static tscratch* scratch_alloc(int threads, int32_t B) {
  tscratch* ts = (tscratch*)calloc((size_t)threads, sizeof(tscratch));
  for (int t = 0; t < threads; ++t) {
    ts[t].Rr = (int64_t*)malloc(sizeof(int64_t) * (size_t)B);
    ts[t].Rs = (int64_t*)malloc(sizeof(int64_t) * (size_t)B);
    ts[t].Cr = (int64_t*)malloc(sizeof(int64_t) * (size_t)B);
    ts[t].Cs = (int64_t*)malloc(sizeof(int64_t) * (size_t)B);
    ts[t].prob = (int64_t*)malloc(sizeof(int64_t) * (size_t)B);
  }
  return ts;
}

static void scratch_free(tscratch* ts, int threads) {
  for (int t = 0; t < threads; ++t) {
    free(ts[t].Rr);
    free(ts[t].Rs);
    free(ts[t].Cr);
    free(ts[t].Cs);
    free(ts[t].prob);
  }
  free(ts);
}

/* The state's scratch, reused by every call and allocated again only for
 * more threads or a larger B.  Buffers allocated per call would make a
 * call's time depend on the allocator (the page faults of freshly mapped
 * memory when the heap has been trimmed) rather than on the algorithm.
 * Nothing reads scratch it has not written first (every proposal clears the
 * rows it uses), so reuse leaves every result as it was. */
static tscratch* scratch_get(orc_state* st, int threads) {
  if (st->ts && st->ts_threads >= threads && st->ts_B >= st->B) return st->ts;
  if (st->ts) scratch_free(st->ts, st->ts_threads);
  const int t = threads > st->p.threads ? threads : st->p.threads;
  st->ts = scratch_alloc(t, st->B);
  st->ts_threads = t;
  st->ts_B = st->B;
  return st->ts;
}

/* ------------------------------------------------------------------------ */
/* State, M build                                                            */
/* ------------------------------------------------------------------------ */

void orc_default_params(orc_params* p) {
  p->K = 10;
  p->reduction_rate = 0.5;
  p->max_nodal_itr = 100;
  p->num_batch_nodal_update = 4;
  p->nodal_batch = 0;
  p->thr1 = 5e-4;
  p->thr2 = 1e-4;
  p->avg_window = 3;
  p->dfs_depth = 20;
  p->block_size = 1024;
  p->init = 0;
  p->ds_mode = 0;
  p->seed = 0;
  p->threads = 16;
}

static inline int64_t wout(const orc_graph* g, int64_t e) { return g->out_w ? g->out_w[e] : 1; }
static inline int64_t win(const orc_graph* g, int64_t e) { return g->in_w ? g->in_w[e] : 1; }

orc_state* orc_create(const orc_graph* g, const orc_params* p) {
  orc_state* st = (orc_state*)calloc(1, sizeof(orc_state));
  st->g = *g;
  st->p = *p;
  if (st->p.threads < 1) st->p.threads = 1;
  int64_t wtot = 0;
  for (int64_t e = 0; e < g->E; ++e) wtot += wout(g, e);
  st->lam_n = wtot + 1;
  st->lam = (int64_t*)malloc(sizeof(int64_t) * (size_t)st->lam_n);
  orc_lambda_table(wtot, st->lam);
  st->labels = (int32_t*)calloc((size_t)g->N, sizeof(int32_t));
  st->row_idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)(g->E + 1));
  st->row_w = (int64_t*)malloc(sizeof(int64_t) * (size_t)(g->E + 1));
  st->col_idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)(g->E + 1));
  st->col_w = (int64_t*)malloc(sizeof(int64_t) * (size_t)(g->E + 1));
  return st;
}

static void free_blocks(orc_state* st) {
  free(st->M);
  free(st->row_ptr);
  free(st->col_ptr);
  free(st->d_out);
  free(st->d_in);
  free(st->d_a);
  free(st->best_s);
  free(st->best_dS);
  st->M = NULL;
  st->row_ptr = st->col_ptr = NULL;
  st->d_out = st->d_in = st->d_a = NULL;
  st->best_s = NULL;
  st->best_dS = NULL;
}

void orc_destroy(orc_state* st) {
  if (!st) return;
  free_blocks(st);
  free(st->lam);
  free(st->labels);
  free(st->row_idx);
  free(st->row_w);
  free(st->col_idx);
  free(st->col_w);
  free(st->plog);
  free(st->uni);
  if (st->ts) scratch_free(st->ts, st->ts_threads);
  free(st);
}

int orc_set_uniforms(orc_state* st, const double* u, int64_t n) {
  if (n < 0 || n % 4 || (n && !u)) return -1;
  for (int64_t i = 0; i < n; ++i)
    if (!(u[i] >= 0.0 && u[i] < 1.0)) return -1;
  free(st->uni);
  st->uni = NULL;
  st->uni_n = st->uni_pos = 0;
  if (n) {
    st->uni = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) st->uni[i] = (uint64_t)ldexp(u[i], 64); /* floor(u 2^64) */
    st->uni_n = n / 4;
  }
  return 0;
}

int64_t orc_uniforms_used(const orc_state* st) { return st->uni_pos; }

static inline int is_dense(const orc_state* st) { return st->B < st->p.block_size; }

/* _initialize_edge_counts, graph_parallel.hpp:703-759 */
static void build_M(orc_state* st) {
  const orc_graph* g = &st->g;
  int32_t B = st->B;
  if (B > st->cap_B) {
    free_blocks(st);
    st->cap_B = B;
    st->d_out = (int64_t*)malloc(sizeof(int64_t) * (size_t)B);
    st->d_in = (int64_t*)malloc(sizeof(int64_t) * (size_t)B);
    st->d_a = (int64_t*)malloc(sizeof(int64_t) * (size_t)B);
    st->row_ptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(B + 1));
    st->col_ptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(B + 1));
    st->best_s = (int32_t*)malloc(sizeof(int32_t) * (size_t)B);
    st->best_dS = (i128*)malloc(sizeof(i128) * (size_t)B);
  }
  free(st->M);
  st->M = NULL;
  memset(st->d_out, 0, sizeof(int64_t) * (size_t)B);
  memset(st->d_in, 0, sizeof(int64_t) * (size_t)B);
  if (is_dense(st)) {
    st->M = (int64_t*)calloc((size_t)B * (size_t)B, sizeof(int64_t));
    for (int32_t x = 0; x < g->N; ++x) {
      int64_t k1 = st->labels[x];
      for (int64_t e = g->out_ptr[x]; e < g->out_ptr[x + 1]; ++e)
        st->M[k1 * B + st->labels[g->out_idx[e]]] += wout(g, e);
    }
    for (int32_t i = 0; i < B; ++i)
      for (int32_t j = 0; j < B; ++j) {
        st->d_out[i] += st->M[(int64_t)i * B + j];
        st->d_in[i] += st->M[(int64_t)j * B + i];
      }
  } else {
    memset(st->row_ptr, 0, sizeof(int64_t) * (size_t)(B + 1));
    memset(st->col_ptr, 0, sizeof(int64_t) * (size_t)(B + 1));
    for (int32_t x = 0; x < g->N; ++x)
      for (int64_t e = g->out_ptr[x]; e < g->out_ptr[x + 1]; ++e) {
        st->row_ptr[st->labels[x] + 1]++;
        st->col_ptr[st->labels[g->out_idx[e]] + 1]++;
      }
    for (int32_t b = 0; b < B; ++b) {
      st->row_ptr[b + 1] += st->row_ptr[b];
      st->col_ptr[b + 1] += st->col_ptr[b];
    }
    int64_t* rfill = (int64_t*)malloc(sizeof(int64_t) * (size_t)B);
    int64_t* cfill = (int64_t*)malloc(sizeof(int64_t) * (size_t)B);
    memcpy(rfill, st->row_ptr, sizeof(int64_t) * (size_t)B);
    memcpy(cfill, st->col_ptr, sizeof(int64_t) * (size_t)B);
    for (int32_t x = 0; x < g->N; ++x) {
      int32_t k1 = st->labels[x];
      for (int64_t e = g->out_ptr[x]; e < g->out_ptr[x + 1]; ++e) {
        int32_t k2 = st->labels[g->out_idx[e]];
        int64_t w = wout(g, e);
        st->row_idx[rfill[k1]] = k2;
        st->row_w[rfill[k1]++] = w;
        st->col_idx[cfill[k2]] = k1;
        st->col_w[cfill[k2]++] = w;
        st->d_out[k1] += w;
        st->d_in[k2] += w;
      }
    }
    free(rfill);
    free(cfill);
  }
  for (int32_t b = 0; b < B; ++b) st->d_a[b] = st->d_out[b] + st->d_in[b];
}

int orc_set_labels(orc_state* st, const int32_t* labels, int32_t B) {
  if (B < 1) return -1;
  for (int32_t x = 0; x < st->g.N; ++x)
    if (labels[x] < 0 || labels[x] >= B) return -1;
  memcpy(st->labels, labels, sizeof(int32_t) * (size_t)st->g.N);
  st->B = B;
  build_M(st);
  return 0;
}

int32_t orc_num_blocks(const orc_state* st) { return st->B; }

void orc_get_labels(const orc_state* st, int32_t* out) {
  memcpy(out, st->labels, sizeof(int32_t) * (size_t)st->g.N);
}

void orc_get_degrees(const orc_state* st, int64_t* d_out, int64_t* d_in) {
  memcpy(d_out, st->d_out, sizeof(int64_t) * (size_t)st->B);
  memcpy(d_in, st->d_in, sizeof(int64_t) * (size_t)st->B);
}

int orc_get_M_dense(const orc_state* st, int64_t* out) {
  int32_t B = st->B;
  if (is_dense(st)) {
    memcpy(out, st->M, sizeof(int64_t) * (size_t)B * (size_t)B);
    return 0;
  }
  memset(out, 0, sizeof(int64_t) * (size_t)B * (size_t)B);
  for (int32_t a = 0; a < B; ++a)
    for (int64_t e = st->row_ptr[a]; e < st->row_ptr[a + 1]; ++e)
      out[(int64_t)a * B + st->row_idx[e]] += st->row_w[e];
  return 0;
}

/* dense copies of row a and column a into scratch (NewM fill,
 * graph_parallel.hpp:1073-1094) */
static void fill_row(const orc_state* st, int32_t a, int64_t* R) {
  int32_t B = st->B;
  if (is_dense(st)) {
    memcpy(R, st->M + (int64_t)a * B, sizeof(int64_t) * (size_t)B);
  } else {
    memset(R, 0, sizeof(int64_t) * (size_t)B);
    for (int64_t e = st->row_ptr[a]; e < st->row_ptr[a + 1]; ++e) R[st->row_idx[e]] += st->row_w[e];
  }
}

static void fill_col(const orc_state* st, int32_t a, int64_t* C) {
  int32_t B = st->B;
  if (is_dense(st)) {
    for (int32_t i = 0; i < B; ++i) C[i] = st->M[(int64_t)i * B + a];
  } else {
    memset(C, 0, sizeof(int64_t) * (size_t)B);
    for (int64_t e = st->col_ptr[a]; e < st->col_ptr[a + 1]; ++e) C[st->col_idx[e]] += st->col_w[e];
  }
}

/* ------------------------------------------------------------------------ */
/* Entropy, graph_parallel.hpp:1281-1319                                     */
/* ------------------------------------------------------------------------ */

double orc_model_S(int64_t N, int64_t E, int64_t B) {
  double x = (double)B * (double)B / (double)E;
  return (double)E * (1.0 + x) * log(1.0 + x) - x * log(x) + (double)N * log((double)B);
}

static i128 data_fx(const orc_state* st) {
  int32_t B = st->B;
  dsacc a = {0, 0};
  int64_t* R = (int64_t*)malloc(sizeof(int64_t) * (size_t)B);
  for (int32_t i = 0; i < B; ++i) {
    fill_row(st, i, R);
    for (int32_t j = 0; j < B; ++j) acc_T(st, &a, -1, R[j], st->d_out[i], st->d_in[j]);
  }
  free(R);
  return a.fx;
}

void orc_entropy(const orc_state* st, int64_t* data_hi, uint64_t* data_lo, double* S) {
  i128 d = data_fx(st);
  if (data_hi) split128(d, data_hi, data_lo);
  if (S) *S = orc_model_S(st->g.N, st->g.E, st->B) + fx_to_double(d);
}

double orc_data_S_double(const orc_state* st) {
  int32_t B = st->B;
  double s = 0;
  int64_t* R = (int64_t*)malloc(sizeof(int64_t) * (size_t)B);
  for (int32_t i = 0; i < B; ++i) {
    fill_row(st, i, R);
    for (int32_t j = 0; j < B; ++j)
      if (R[j] != 0)
        s -= (double)R[j] * log((double)R[j] / ((double)st->d_out[i] * (double)st->d_in[j]));
  }
  free(R);
  return s;
}

/* ------------------------------------------------------------------------ */
/* Block merge, graph_parallel.hpp:761-988 and :349-368                      */
/* ------------------------------------------------------------------------ */

static void merge_propose(orc_state* st, tscratch* ts, int32_t r, int32_t idx, uint64_t stream,
                          int32_t* s_out, i128* dS_out) {
  int32_t B = st->B;
  int64_t *Rr = ts->Rr, *Rs = ts->Rs, *Cr = ts->Cr, *Cs = ts->Cs, *prob = ts->prob;
  const int64_t *dout = st->d_out, *din = st->d_in, *da = st->d_a;
  fill_row(st, r, Rr);
  fill_col(st, r, Cr);
  int64_t k_out = 0, k_in = 0;
  for (int32_t i = 0; i < B; ++i) {
    k_out += Rr[i];
    k_in += Cr[i];
    prob[i] = Rr[i] + Cr[i];
  }
  int64_t k = k_out + k_in;
  uint64_t x[4];
  if (st->uni) /* pre-drawn: stream = the round's first draw */
    pre_draw(st, stream + (uint64_t)r * (uint64_t)st->p.K + (uint64_t)idx, x);
  else
    draw4(st->p.seed, (uint64_t)r, (uint64_t)idx, stream, ORC_TAG_MERGE, x);
  int32_t s;
  if (k == 0) {
    s = (int32_t)idx_of(x[2], B);
  } else {
    int32_t u = (int32_t)orc_pick(prob, B, idx_of(x[0], k));
    if (uniform_branch(x[1], da[u], B)) {
      s = uniform_not_r(x[2], r, B);
    } else {
      fill_row(st, u, Rs); /* reuse Rs/Cs as the multinomial scratch */
      fill_col(st, u, Cs);
      int64_t W = 0;
      for (int32_t i = 0; i < B; ++i) {
        prob[i] = (i == r) ? 0 : Rs[i] + Cs[i];
        W += prob[i];
      }
      if (W == 0)
        s = uniform_not_r(x[2], r, B);
      else
        s = (int32_t)orc_pick(prob, B, idx_of(x[3], W));
    }
  }
  *s_out = s;
  if (s == r) { /* only reachable with k == 0 or B == 1: dS is 0 */
    *dS_out = 0;
    return;
  }
  fill_row(st, s, Rs);
  fill_col(st, s, Cs);
  int64_t dout_new_s = dout[s] + k_out;
  int64_t din_new_s = din[s] + k_in;
  dsacc a = {0, 0};
  acc_T(st, &a, +1, Rr[r], din[r], dout[r]);
  acc_T(st, &a, +1, Rr[s], din[s], dout[r]);
  acc_T(st, &a, +1, Rs[r], din[r], dout[s]);
  acc_T(st, &a, +1, Rs[s], din[s], dout[s]);
  acc_T(st, &a, -1, Rr[r] + Rr[s] + Rs[r] + Rs[s], din_new_s, dout_new_s);
  for (int32_t i = 0; i < B; ++i) {
    if (i == r || i == s) continue;
    acc_T(st, &a, +1, Rr[i], din[i], dout[r]);
    acc_T(st, &a, +1, Rs[i], din[i], dout[s]);
    acc_T(st, &a, -1, Rr[i] + Rs[i], din[i], dout_new_s);
    acc_T(st, &a, +1, Cr[i], dout[i], din[r]);
    acc_T(st, &a, +1, Cs[i], dout[i], din[s]);
    acc_T(st, &a, -1, Cr[i] + Cs[i], dout[i], din_new_s);
  }
  if (a.inf != 0) abort(); /* merge dS is always finite */
  *dS_out = a.fx;
}

void orc_merge_propose_one(orc_state* st, int32_t r, int32_t idx, uint64_t stream, int32_t* s,
                           int64_t* dS_hi, uint64_t* dS_lo) {
  if (st->uni) {
    if ((int64_t)st->B * st->p.K > uni_left(st)) {
      *s = -1;
      return;
    }
    stream = (uint64_t)st->uni_pos;
  }
  tscratch* ts = scratch_get(st, 1);
  i128 d;
  merge_propose(st, ts, r, idx, stream, s, &d);
  split128(d, dS_hi, dS_lo);
}

typedef struct {
  orc_state* st;
  tscratch* ts;
  uint64_t stream;
  int64_t first; /* block of item 0 */
} merge_ctx;

static void merge_item(void* vp, int tid, int64_t i) {
  merge_ctx* c = (merge_ctx*)vp;
  orc_state* st = c->st;
  const int64_t r = c->first + i;
  int32_t best = -1;
  i128 bestd = 0;
  for (int32_t idx = 0; idx < st->p.K; ++idx) {
    int32_t s;
    i128 d;
    merge_propose(st, &c->ts[tid], (int32_t)r, idx, c->stream, &s, &d);
    if (best < 0 || d < bestd) { /* strict less: the first proposal wins ties */
      best = s;
      bestd = d;
    }
  }
  st->best_s[r] = best;
  st->best_dS[r] = bestd;
}

int orc_block_merge(orc_state* st, uint64_t stream, int32_t* best_s, int64_t* dS_hi,
                    uint64_t* dS_lo) {
  if (st->uni) {
    if ((int64_t)st->B * st->p.K > uni_left(st)) return -1;
    stream = (uint64_t)st->uni_pos;
    st->uni_pos += (int64_t)st->B * st->p.K;
  }
  merge_ctx c;
  c.st = st;
  c.ts = scratch_get(st, st->p.threads);
  c.stream = stream;
  c.first = 0;
  par_for(st->p.threads, st->B, merge_item, &c);
  for (int32_t r = 0; r < st->B; ++r) {
    if (best_s) best_s[r] = st->best_s[r];
    if (dS_hi) split128(st->best_dS[r], &dS_hi[r], &dS_lo[r]);
  }
  return 0;
}

void orc_merge_eval(orc_state* st, int32_t r0, int32_t r1, uint64_t stream) {
  if (r0 < 0) r0 = 0;
  if (r1 > st->B) r1 = st->B;
  if (r1 <= r0) return;
  if (st->uni) {
    if ((int64_t)st->B * st->p.K > uni_left(st)) return;
    stream = (uint64_t)st->uni_pos;
  }
  merge_ctx c;
  c.st = st;
  c.ts = scratch_get(st, st->p.threads);
  c.stream = stream;
  c.first = r0;
  par_for(st->p.threads, r1 - r0, merge_item, &c);
}

/* ------------------------------------------------------------------------ */
/* Carry out best merges, graph_parallel.hpp:1225-1279                       */
/* ------------------------------------------------------------------------ */

static const orc_state* g_sort_st;
static int cmp_blocks(const void* a, const void* b) {
  int32_t i = *(const int32_t*)a, j = *(const int32_t*)b;
  i128 di = g_sort_st->best_dS[i], dj = g_sort_st->best_dS[j];
  if (di < dj) return -1;
  if (di > dj) return 1;
  return (i > j) - (i < j); /* ties by block id (the reference's sort is unstable) */
}

int orc_carry_out(orc_state* st, int32_t B_to_merge) {
  int32_t B = st->B, N = st->g.N;
  int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)B);
  int32_t* block_map = (int32_t*)malloc(sizeof(int32_t) * (size_t)B);
  int32_t* lab = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
  for (int32_t i = 0; i < B; ++i) order[i] = block_map[i] = i;
  g_sort_st = st;
  qsort(order, (size_t)B, sizeof(int32_t), cmp_blocks);
  memcpy(lab, st->labels, sizeof(int32_t) * (size_t)N);
  int32_t num_merge = 0, counter = 0;
  int rc = 0;
  while (num_merge < B_to_merge) {
    if (counter >= B) {
      rc = -1;
      break;
    }
    int32_t from = order[counter];
    int32_t to = block_map[st->best_s[from]];
    counter++;
    if (to != from) {
      for (int32_t i = 0; i < B; ++i)
        if (block_map[i] == from) block_map[i] = to;
      for (int32_t x = 0; x < N; ++x)
        if (lab[x] == from) lab[x] = to;
      num_merge++;
    }
  }
  if (rc == 0) {
    /* compact: rank among the surviving labels, ascending */
    char* seen = (char*)calloc((size_t)B, 1);
    for (int32_t x = 0; x < N; ++x) seen[lab[x]] = 1;
    int32_t next = 0;
    for (int32_t b = 0; b < B; ++b) block_map[b] = seen[b] ? next++ : -1;
    for (int32_t x = 0; x < N; ++x) lab[x] = block_map[lab[x]];
    free(seen);
    memcpy(st->labels, lab, sizeof(int32_t) * (size_t)N);
    st->B = B - B_to_merge;
    build_M(st);
  }
  free(order);
  free(block_map);
  free(lab);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Nodal proposal, graph_parallel.hpp:990-1222                               */
/* ------------------------------------------------------------------------ */

static void nodal_propose(orc_state* st, tscratch* ts, int32_t n, uint64_t stream, int32_t* s_out,
                          dsacc* dS, int* accept) {
  const orc_graph* g = &st->g;
  int32_t B = st->B;
  const int32_t* lab = st->labels;
  const int64_t *dout = st->d_out, *din = st->d_in, *da = st->d_a;
  int32_t r = lab[n];
  int64_t o0 = g->out_ptr[n], o1 = g->out_ptr[n + 1];
  int64_t i0 = g->in_ptr[n], i1 = g->in_ptr[n + 1];
  int64_t k_out = 0, k_in = 0;
  for (int64_t e = o0; e < o1; ++e) k_out += wout(g, e);
  for (int64_t e = i0; e < i1; ++e) k_in += win(g, e);
  int64_t k = k_out + k_in;
  uint64_t x[4];
  if (st->uni) /* pre-drawn: node n takes draw stream + n */
    pre_draw(st, stream + (uint64_t)n, x);
  else
    draw4(st->p.seed, (uint64_t)n, 0, stream, ORC_TAG_NODAL, x);
  int32_t s;
  if (k == 0) {
    s = (int32_t)idx_of(x[2], B);
  } else {
    /* neighbour pick: out-neighbours then in-neighbours, weight w */
    int64_t t = idx_of(x[0], k), acc = 0;
    int32_t v = -1;
    for (int64_t e = o0; e < o1 && v < 0; ++e) {
      acc += wout(g, e);
      if (acc > t) v = g->out_idx[e];
    }
    for (int64_t e = i0; e < i1 && v < 0; ++e) {
      acc += win(g, e);
      if (acc > t) v = g->in_idx[e];
    }
    int32_t u = lab[v];
    if (uniform_branch(x[1], da[u], B)) {
      s = (int32_t)idx_of(x[2], B);
    } else {
      fill_row(st, u, ts->Rs);
      fill_col(st, u, ts->Cs);
      for (int32_t i = 0; i < B; ++i) ts->prob[i] = ts->Rs[i] + ts->Cs[i];
      s = (int32_t)orc_pick(ts->prob, B, idx_of(x[3], da[u]));
    }
  }
  *s_out = s;
  dS->fx = 0;
  dS->inf = 0;
  *accept = 0;
  if (s == r) return;

  int64_t *Rr = ts->Rr, *Rs = ts->Rs, *Cr = ts->Cr, *Cs = ts->Cs;
  fill_row(st, r, Rr);
  fill_row(st, s, Rs);
  fill_col(st, r, Cr);
  fill_col(st, s, Cs);
  int64_t dout_new_r = dout[r] - k_out, dout_new_s = dout[s] + k_out;
  int64_t din_new_r = din[r] - k_in, din_new_s = din[s] + k_in;
  for (int32_t i = 0; i < B; ++i) {
    acc_T(st, dS, +1, Rr[i], din[i], dout[r]);
    acc_T(st, dS, +1, Rs[i], din[i], dout[s]);
    if (i != r && i != s) {
      acc_T(st, dS, +1, Cr[i], dout[i], din[r]);
      acc_T(st, dS, +1, Cs[i], dout[i], din[s]);
    }
  }
  const int exact = st->p.ds_mode == 1;
  for (int64_t e = o0; e < o1; ++e) {
    int32_t v = g->out_idx[e];
    int64_t w = wout(g, e);
    int32_t b = lab[v];
    if (exact && v == n) { /* self-loop moves (r,r) -> (s,s) */
      Rr[r] -= w;
      Rs[s] += w;
    } else {
      Rr[b] -= w;
      Rs[b] += w;
    }
  }
  for (int64_t e = i0; e < i1; ++e) {
    int32_t v = g->in_idx[e];
    int64_t w = win(g, e);
    int32_t b = lab[v];
    if (exact && v != n) { /* graph.cu:1414-1426 */
      if (b == r) {
        Rr[r] -= w;
        Rr[s] += w;
      }
      if (b == s) {
        Rs[r] -= w;
        Rs[s] += w;
      }
    }
    Cr[b] -= w;
    Cs[b] += w;
  }
  acc_T(st, dS, -1, Rr[r], din_new_r, dout_new_r);
  acc_T(st, dS, -1, Rr[s], din_new_s, dout_new_r);
  acc_T(st, dS, -1, Rs[r], din_new_r, dout_new_s);
  acc_T(st, dS, -1, Rs[s], din_new_s, dout_new_s);
  for (int32_t i = 0; i < B; ++i) {
    if (i == r || i == s) continue;
    acc_T(st, dS, -1, Rr[i], din[i], dout_new_r);
    acc_T(st, dS, -1, Rs[i], din[i], dout_new_s);
    acc_T(st, dS, -1, Cr[i], dout[i], din_new_r);
    acc_T(st, dS, -1, Cs[i], dout[i], din_new_s);
  }
  *accept = (dS->inf == -1) || (dS->inf == 0 && dS->fx < 0);
}

void orc_nodal_propose_one(orc_state* st, int32_t n, uint64_t stream, int32_t* s, int64_t* dS_hi,
                           uint64_t* dS_lo, int32_t* inf, int32_t* accept) {
  if (st->uni) {
    if ((int64_t)n >= uni_left(st)) {
      *s = -1;
      *inf = 2;
      *accept = 0;
      return;
    }
    stream = (uint64_t)st->uni_pos;
  }
  tscratch* ts = scratch_get(st, 1);
  dsacc d;
  int acc;
  nodal_propose(st, ts, n, stream, s, &d, &acc);
  split128(d.fx, dS_hi, dS_lo);
  *inf = d.inf;
  *accept = acc;
}

typedef struct {
  orc_state* st;
  tscratch* ts;
  uint64_t stream;
  int32_t beg;
  int32_t* s;
  dsacc* d;
  int* acc;
} nodal_ctx;

static void nodal_item(void* vp, int tid, int64_t i) {
  nodal_ctx* c = (nodal_ctx*)vp;
  nodal_propose(c->st, &c->ts[tid], c->beg + (int32_t)i, c->stream, &c->s[i], &c->d[i], &c->acc[i]);
}

static void nodal_run(orc_state* st, int32_t beg, int32_t end, uint64_t stream, int32_t* s,
                      dsacc* d, int* acc) {
  nodal_ctx c;
  c.st = st;
  c.ts = scratch_get(st, st->p.threads);
  c.stream = stream;
  c.beg = beg;
  c.s = s;
  c.d = d;
  c.acc = acc;
  par_for(st->p.threads, end - beg, nodal_item, &c);
}

/* nodal_move + update, graph_parallel.hpp:406-458 (without the convergence
 * test, which is in orc_partition) */
int orc_nodal_batch(orc_state* st, int32_t beg, int32_t end, uint64_t stream, int32_t* s_out,
                    int64_t* dS_hi, uint64_t* dS_lo, int32_t* inf_out, int32_t* acc_out,
                    int64_t* sum_hi, uint64_t* sum_lo, int32_t* sum_inf, int32_t* n_moves) {
  int32_t n = end - beg;
  if (st->uni) {
    if ((int64_t)n > uni_left(st)) return -1;
    stream = (uint64_t)(st->uni_pos - beg);
    st->uni_pos += n;
  }
  int32_t* s = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  dsacc* d = (dsacc*)malloc(sizeof(dsacc) * (size_t)(n > 0 ? n : 1));
  int* acc = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  nodal_run(st, beg, end, stream, s, d, acc);
  dsacc sum = {0, 0};
  int32_t moves = 0;
  for (int32_t i = 0; i < n; ++i) {
    if (s_out) s_out[i] = s[i];
    if (dS_hi) split128(d[i].fx, &dS_hi[i], &dS_lo[i]);
    if (inf_out) inf_out[i] = d[i].inf;
    if (acc_out) acc_out[i] = acc[i];
    if (acc[i]) {
      moves++;
      if (d[i].inf)
        acc_inf(&sum, d[i].inf);
      else
        sum.fx += d[i].fx;
    }
  }
  for (int32_t i = 0; i < n; ++i)
    if (acc[i]) st->labels[beg + i] = s[i];
  build_M(st);
  if (sum_hi) split128(sum.fx, sum_hi, sum_lo);
  if (sum_inf) *sum_inf = sum.inf;
  if (n_moves) *n_moves = moves;
  free(s);
  free(d);
  free(acc);
  return 0;
}

void orc_nodal_eval(orc_state* st, int32_t beg, int32_t end, uint64_t stream, int32_t* n_accept) {
  int32_t n = end - beg;
  if (st->uni) {
    if ((int64_t)n > uni_left(st)) return;
    stream = (uint64_t)(st->uni_pos - beg);
  }
  int32_t* s = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  dsacc* d = (dsacc*)malloc(sizeof(dsacc) * (size_t)(n > 0 ? n : 1));
  int* acc = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  nodal_run(st, beg, end, stream, s, d, acc);
  int32_t a = 0;
  for (int32_t i = 0; i < n; ++i) a += acc[i];
  if (n_accept) *n_accept = a;
  free(s);
  free(d);
  free(acc);
}

/* ------------------------------------------------------------------------ */
/* SCC initial partition, graph_parallel.hpp:514-603                         */
/* ------------------------------------------------------------------------ */

int32_t orc_scc_init(const orc_graph* g, int32_t dfs_depth, int32_t* labels) {
  int32_t N = g->N;
  int64_t E = g->E;
  /* _dfs_itr: iterative DFS over out-edges; finish order onto a stack */
  int32_t* color = (int32_t*)calloc((size_t)N, sizeof(int32_t));
  int32_t* finish = (int32_t*)malloc(sizeof(int32_t) * (size_t)(N > 0 ? N : 1));
  int64_t nfin = 0;
  int32_t* stk = (int32_t*)malloc(sizeof(int32_t) * (size_t)(N + E + 1));
  for (int32_t v0 = 0; v0 < N; ++v0) {
    if (color[v0] != 0) continue;
    int64_t sp = 0;
    stk[sp++] = v0;
    while (sp > 0) {
      int32_t v = stk[sp - 1];
      if (color[v] != 0) {
        sp--;
        if (color[v] == 1) {
          finish[nfin++] = v;
          color[v] = 2;
        }
      } else {
        color[v] = 1;
        for (int64_t e = g->out_ptr[v]; e < g->out_ptr[v + 1]; ++e)
          if (color[g->out_idx[e]] == 0) stk[sp++] = g->out_idx[e];
      }
    }
  }
  /* transpose, built by scanning sources ascending (graph_parallel.hpp:517-522) */
  int64_t* tp = (int64_t*)calloc((size_t)N + 1, sizeof(int64_t));
  int32_t* ti = (int32_t*)malloc(sizeof(int32_t) * (size_t)(E + 1));
  for (int64_t e = 0; e < E; ++e) tp[g->out_idx[e] + 1]++;
  for (int32_t v = 0; v < N; ++v) tp[v + 1] += tp[v];
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(N + 1));
  memcpy(fill, tp, sizeof(int64_t) * (size_t)(N + 1));
  for (int32_t i = 0; i < N; ++i)
    for (int64_t e = g->out_ptr[i]; e < g->out_ptr[i + 1]; ++e) ti[fill[g->out_idx[e]]++] = i;
  /* depth-limited DFS on the transpose in reverse finish order */
  char* visited = (char*)calloc((size_t)N, 1);
  int32_t b = 0;
  while (nfin > 0) {
    int32_t v = finish[--nfin];
    if (visited[v]) continue;
    int32_t cnt = 0;
    int64_t sp = 0;
    stk[sp++] = v;
    while (sp > 0 && cnt < dfs_depth) {
      int32_t ni = stk[--sp];
      if (!visited[ni]) {
        visited[ni] = 1;
        labels[ni] = b;
        cnt++;
        for (int64_t e = tp[ni]; e < tp[ni + 1]; ++e)
          if (!visited[ti[e]]) stk[sp++] = ti[e];
      }
    }
    b++;
  }
  free(color);
  free(finish);
  free(stk);
  free(tp);
  free(ti);
  free(fill);
  free(visited);
  return b;
}

/* ------------------------------------------------------------------------ */
/* Golden section, graph_parallel.hpp:1322-1365                              */
/* ------------------------------------------------------------------------ */

int orc_golden_update(orc_snap slots[3], orc_snap cur, double rate, orc_snap* next,
                      int64_t* B_to_merge) {
  orc_snap *large = &slots[0], *med = &slots[1], *small = &slots[2];
  if (cur.S <= med->S) {
    if (med->B > cur.B)
      *large = *med;
    else
      *small = *med;
    *med = cur;
  } else {
    if (med->B > cur.B)
      *small = cur;
    else
      *large = cur;
  }
  *B_to_merge = 0;
  if (isinf(small->S)) {
    *next = *med;
    *B_to_merge = (int64_t)floor((double)med->B * rate);
    return *B_to_merge == 0 ? 1 : 0;
  }
  if (large->id < 0 || large->B - small->B == 2) { /* large unset: stop at med */
    *next = *med;
    return 1;
  }
  if ((large->B - med->B) >= (med->B - small->B)) {
    int64_t nb = med->B + (int64_t)round((double)(large->B - med->B) * 0.618);
    *next = *large;
    *B_to_merge = large->B - nb;
  } else {
    int64_t nb = small->B + (int64_t)round((double)(med->B - small->B) * 0.618);
    *next = *med;
    *B_to_merge = med->B - nb;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* The nodal phase, graph_parallel.hpp:381-478                               */
/* ------------------------------------------------------------------------ */

/* What one nodal phase needs from the partition state: the MDL at a sweep's
 * start (prepare_batch, :383) and one batch (nodal_move + update, :406-458:
 * proposals against the frozen state, then the accepted moves applied),
 * returning the exact sum of the batch's accepted dS and its move count. */
typedef struct {
  void* ctx;
  double (*entropy)(void* ctx);
  void (*batch)(void* ctx, int32_t beg, int32_t end, uint64_t stream, dsacc* sum,
                int32_t* moves);
} nodal_ops;

typedef struct {
  int32_t sweeps;
  int64_t batches, proposals, moves;
  int converged;
  double S_last; /* S after the last executed batch */
} nodal_result;

/* Sweeps of contiguous batches [beg, beg + batch) until the moving-average
 * test fires or max_nodal_itr sweeps (graph_parallel.hpp:381-478):
 *   prepare_batch :381-395  S = MDL; mean of itr_delta_entropy over the
 *                           avg_window - 1 sweeps before this one (itr >=
 *                           window - 1); the sweep's dS accumulator zeroed
 *   update        :433-458  dS_sum = the sweep's accumulator, cumulative over
 *                           its batches so far; itr_delta_entropy[itr] =
 *                           dS_sum; S += dS_sum (cumulative);
 *                           converged when itr >= window - 1 and
 *                           -mean < thr * S, thr = thr2 if isf else thr1
 *   check_batch   :460-468  next batch while beg < N
 *   check_itr     :470-478  next sweep while itr < max_nodal_itr
 * Batch j of the phase uses random stream *stream + j; *stream advances by
 * the batches run.  itr_dS has max_nodal_itr entries. */
static void nodal_phase(const orc_params* p, int32_t N, int32_t batch, int isf,
                        const nodal_ops* ops, uint64_t* stream, double* itr_dS,
                        nodal_result* res) {
  const int32_t W = p->avg_window;
  memset(res, 0, sizeof(*res));
  memset(itr_dS, 0, sizeof(double) * (size_t)p->max_nodal_itr);
  for (int32_t itr = 0; itr < p->max_nodal_itr; ++itr) {
    double S = ops->entropy(ops->ctx);
    double mean = 0;
    if (itr >= W - 1) {
      for (int32_t i = itr - W + 1; i < itr; ++i) mean += itr_dS[i];
      mean /= (double)(W - 1);
    }
    dsacc cum = {0, 0};
    res->sweeps += 1;
    for (int32_t beg = 0; beg < N; beg += batch) {
      const int32_t end = beg + batch > N ? N : beg + batch;
      dsacc sum = {0, 0};
      int32_t moves = 0;
      ops->batch(ops->ctx, beg, end, (*stream)++, &sum, &moves);
      res->batches += 1;
      res->proposals += end - beg;
      res->moves += moves;
      if (sum.inf)
        acc_inf(&cum, sum.inf);
      else
        cum.fx += sum.fx;
      const double dS_sum = acc_to_double(&cum);
      itr_dS[itr] = dS_sum;
      S += dS_sum;
      res->S_last = S;
      if (itr >= W - 1) {
        const double thr = isf ? p->thr2 : p->thr1;
        if (-mean < thr * S) {
          res->converged = 1;
          return;
        }
      }
    }
  }
}

static double state_entropy(void* ctx) {
  double S;
  orc_entropy((const orc_state*)ctx, NULL, NULL, &S);
  return S;
}

static void state_batch(void* ctx, int32_t beg, int32_t end, uint64_t stream, dsacc* sum,
                        int32_t* moves) {
  int64_t hi = 0;
  uint64_t lo = 0;
  int32_t sinf = 0;
  *moves = 0;
  /* a pre-drawn stream that runs out leaves the state unchanged; the
   * partition checks the cursor after the phase */
  (void)orc_nodal_batch((orc_state*)ctx, beg, end, stream, NULL, NULL, NULL, NULL, NULL, &hi,
                        &lo, &sinf, moves);
  sum->inf = sinf;
  sum->fx = ((i128)hi << 64) | (i128)lo;
}

typedef struct {
  const double* S_start;
  int64_t n_S;
  const int64_t* hi;
  const uint64_t* lo;
  const int32_t* inf;
  const int32_t* moves;
  int64_t n_batch;
  int64_t sweep;
  int bad;
} script_ctx;

static double script_entropy(void* vp) {
  script_ctx* c = (script_ctx*)vp;
  if (c->sweep >= c->n_S) {
    c->bad = 1;
    return 0.0;
  }
  return c->S_start[c->sweep++];
}

static void script_batch(void* vp, int32_t beg, int32_t end, uint64_t stream, dsacc* sum,
                         int32_t* moves) {
  script_ctx* c = (script_ctx*)vp;
  (void)beg;
  (void)end;
  if ((int64_t)stream >= c->n_batch) {
    c->bad = 1;
    sum->fx = 0;
    sum->inf = 0;
    *moves = 0;
    return;
  }
  sum->fx = ((i128)c->hi[stream] << 64) | (i128)c->lo[stream];
  sum->inf = c->inf[stream];
  *moves = c->moves ? c->moves[stream] : 0;
}

int32_t orc_partition_log(const orc_state* st, int32_t cap, orc_phase* out) {
  for (int32_t i = 0; i < st->plog_n && i < cap && out; ++i) out[i] = st->plog[i];
  return st->plog_n;
}

int orc_nodal_phase_scripted(const orc_params* p, int32_t N, int32_t batch, int32_t isf,
                             const double* S_start, int64_t n_S, const int64_t* sum_hi,
                             const uint64_t* sum_lo, const int32_t* sum_inf,
                             const int32_t* moves, int64_t n_batch, double* out,
                             double* itr_dS) {
  script_ctx c = {S_start, n_S, sum_hi, sum_lo, sum_inf, moves, n_batch, 0, 0};
  nodal_ops ops = {&c, script_entropy, script_batch};
  nodal_result r;
  uint64_t stream = 0;
  double* dS = (double*)calloc((size_t)p->max_nodal_itr, sizeof(double));
  nodal_phase(p, N, batch, isf, &ops, &stream, dS, &r);
  if (itr_dS) memcpy(itr_dS, dS, sizeof(double) * (size_t)p->max_nodal_itr);
  free(dS);
  out[0] = r.sweeps;
  out[1] = (double)r.batches;
  out[2] = (double)r.proposals;
  out[3] = (double)r.moves;
  out[4] = r.converged;
  out[5] = r.S_last;
  return c.bad ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* partition(), graph_parallel.hpp:322-512                                   */
/* ------------------------------------------------------------------------ */

#define ORC_MAX_SAVED 4096

int32_t orc_partition(orc_state* st, int32_t* labels_out, double* stats) {
  const orc_params* p = &st->p;
  int32_t N = st->g.N;
  int32_t batch = p->nodal_batch > 0 ? p->nodal_batch : N / p->num_batch_nodal_update;
  if (batch < 1) batch = 1;
  double* itr_dS = (double*)calloc((size_t)p->max_nodal_itr + 1, sizeof(double));
  int32_t** saved = (int32_t**)calloc(ORC_MAX_SAVED, sizeof(int32_t*));
  int64_t Btm = (int64_t)floor((double)st->B * p->reduction_rate);
  orc_snap slots[3];
  for (int i = 0; i < 3; ++i) {
    slots[i].B = 0;
    slots[i].S = INFINITY;
    slots[i].id = -1;
  }
  double nphase = 0, nsweep = 0, nbatch = 0, nprop = 0, nmoves = 0, nmerge = 0, err = 0;
  st->plog_n = 0;
  uint64_t merge_stream = 0, nodal_stream = 0;
  double S = 0;
  int32_t result = -1;
  for (int64_t phase = 0;; ++phase) {
    /* prepare_merge + block_merge + perform_merge + prepare_move */
    nmerge += (double)st->B * p->K;
    if (orc_block_merge(st, merge_stream++, NULL, NULL, NULL) < 0) {
      err = -4; /* pre-drawn stream exhausted */
      break;
    }
    if (orc_carry_out(st, (int32_t)Btm) < 0) {
      err = -5;
      break;
    }
    nphase += 1;
    /* the nodal phase (prepare_move :374-379, then :381-478); isf is the
     * golden-section bracket's state, fixed during the phase */
    const int isf = isfinite(slots[0].S) && isfinite(slots[1].S) && isfinite(slots[2].S);
    nodal_ops ops = {st, state_entropy, state_batch};
    nodal_result nr;
    const int64_t pos0 = st->uni_pos;
    nodal_phase(p, N, batch, isf, &ops, &nodal_stream, itr_dS, &nr);
    if (st->uni && st->uni_pos - pos0 != nr.proposals) {
      err = -4; /* pre-drawn stream exhausted */
      break;
    }
    nsweep += nr.sweeps;
    nbatch += (double)nr.batches;
    nprop += (double)nr.proposals;
    nmoves += (double)nr.moves;
    /* prepare_for_next */
    orc_entropy(st, NULL, NULL, &S);
    if (st->plog_n == st->plog_cap) {
      st->plog_cap = st->plog_cap ? 2 * st->plog_cap : 64;
      st->plog = (orc_phase*)realloc(st->plog, sizeof(orc_phase) * (size_t)st->plog_cap);
    }
    {
      orc_phase* ph = &st->plog[st->plog_n++];
      ph->B = st->B;
      ph->S = S;
      ph->sweeps = nr.sweeps;
      ph->batches = nr.batches;
      ph->moves = nr.moves;
    }
    if (phase >= ORC_MAX_SAVED) {
      err = -6;
      break;
    }
    saved[phase] = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
    memcpy(saved[phase], st->labels, sizeof(int32_t) * (size_t)N);
    orc_snap cur = {st->B, S, phase};
    orc_snap next;
    int rc = orc_golden_update(slots, cur, p->reduction_rate, &next, &Btm);
    if (next.id != phase) orc_set_labels(st, saved[next.id], (int32_t)next.B);
    for (int64_t i = 0; i <= phase; ++i) {
      if (!saved[i]) continue;
      if (i == slots[0].id || i == slots[1].id || i == slots[2].id || i == next.id) continue;
      free(saved[i]);
      saved[i] = NULL;
    }
    if (rc == 1) {
      result = st->B;
      S = next.S;
      break;
    }
  }
  if (labels_out && result > 0) orc_get_labels(st, labels_out);
  if (stats) {
    stats[0] = S;
    stats[1] = nphase;
    stats[2] = nsweep;
    stats[3] = nbatch;
    stats[4] = nprop;
    stats[5] = nmoves;
    stats[6] = nmerge;
    stats[7] = err;
  }
  for (int64_t i = 0; i < ORC_MAX_SAVED; ++i) free(saved[i]);
  free(saved);
  free(itr_dS);
  return result > 0 ? result : (int32_t)err;
}
