// host.cpp -- host-side utilities of the engine: lambda table, model entropy,
// SCC initial partition, CSR/CSC build, GraphChallenge TSV loader and the
// seeded directed DCSBM generator.  No device code.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "usap.h"
#include "usap_common.hpp"

namespace {

// Sequential draws from the Philox stream (key {seed, tag}, counter {i, stream}).
class Stream {
 public:
  Stream(uint64_t seed, uint64_t stream) : seed_(seed), stream_(stream) {}
  uint64_t u64() {
    if (pos_ == 4) {
      usap::philox4x64(ctr_++, stream_, 0, 0, seed_, 0x5EEDDC5Bull, buf_);
      pos_ = 0;
    }
    return buf_[pos_++];
  }
  double uniform() { return (double)(u64() >> 11) * 0x1.0p-53; }      // [0,1)
  double uniform_pos() { return ((double)(u64() >> 11) + 0.5) * 0x1.0p-53; }  // (0,1)
  int64_t index(int64_t n) { return usap::idx_of(u64(), n); }
  double normal() {  // Box-Muller
    const double u1 = uniform_pos(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  // Gamma(alpha, 1): Marsaglia & Tsang (2000), boosted for alpha < 1.
  double gamma(double alpha) {
    if (alpha < 1.0) return gamma(alpha + 1.0) * std::pow(uniform_pos(), 1.0 / alpha);
    const double d = alpha - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
    for (;;) {
      double x, v;
      do {
        x = normal();
        v = 1.0 + c * x;
      } while (v <= 0);
      v = v * v * v;
      const double u = uniform_pos();
      if (std::log(u) < 0.5 * x * x + d - d * v + d * std::log(v)) return d * v;
    }
  }

 private:
  uint64_t seed_, stream_, ctr_ = 0;
  uint64_t buf_[4] = {0, 0, 0, 0};
  int pos_ = 4;
};

}  // namespace

extern "C" {

void usap_philox4x64(const uint64_t* ctr, const uint64_t* key, uint64_t* out) {
  usap::philox4x64(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1], out);
}

int usap_lambda_table(int64_t maxv, int64_t* out) {
  if (maxv < 0 || !out) return USAP_EINVAL;
  out[0] = 0;
  for (int64_t x = 1; x <= maxv; ++x) out[x] = llrint(std::log((double)x) * 281474976710656.0);
  return USAP_OK;
}

// graph_parallel.hpp:1314-1316: E(1+x)log(1+x) - x log x + N log B, x = B^2/E
double usap_model_S(int64_t N, int64_t E, int64_t B) {
  const double x = (double)B * (double)B / (double)E;
  return (double)E * (1.0 + x) * std::log(1.0 + x) - x * std::log(x) +
         (double)N * std::log((double)B);
}

// _scc + _dfs_itr, graph_parallel.hpp:514-603: finish order of an iterative
// DFS over out-edges, then depth-limited DFS on the transpose in reverse
// finish order; each DFS labels at most dfs_depth nodes as one block.
int usap_scc_init(const usap_graph* g, int32_t dfs_depth, int32_t* labels, int32_t* B) {
  if (!g || !labels || !B || dfs_depth < 1) return USAP_EINVAL;
  const int32_t N = g->N;
  const int64_t E = g->E;
  // The reference's first DFS pushes every unvisited out-neighbour of the
  // node it expands, explores the last one pushed first, skips entries found
  // visited when they come up, and finishes a node when it is on top again.
  // A cursor per stacked node that walks its out-list backwards and descends
  // into the first neighbour still unvisited gives the same expansion and
  // finish order (a neighbour's colour only grows between the push and the
  // pop) with at most N frames instead of one entry per edge.
  std::vector<uint8_t> color((size_t)N, 0);
  std::vector<int32_t> finish;
  finish.reserve((size_t)N);
  std::vector<int32_t> fv((size_t)N);
  std::vector<int64_t> fe((size_t)N);
  for (int32_t root = 0; root < N; ++root) {
    if (color[root]) continue;
    int32_t top = 0;
    fv[0] = root;
    fe[0] = g->out_ptr[root + 1];
    color[root] = 1;
    while (top >= 0) {
      const int32_t v = fv[top];
      const int64_t lo = g->out_ptr[v];
      int64_t e = fe[top];
      int32_t next = -1;
      while (e > lo) {
        const int32_t w = g->out_idx[--e];
        if (!color[w]) {
          next = w;
          break;
        }
      }
      fe[top] = e;
      if (next >= 0) {
        color[next] = 1;
        ++top;
        fv[top] = next;
        fe[top] = g->out_ptr[next + 1];
      } else {
        finish.push_back(v);
        color[v] = 2;
        --top;
      }
    }
  }
  // The transpose listed by source ascending (graph_parallel.hpp:517-522).
  // The in-lists are that transpose whenever each lists its sources in
  // ascending order (edges given by source, as loaded files and the
  // generator give them); otherwise it is rebuilt by a counting sort.
  std::vector<int64_t> tp((size_t)N + 1, 0);
  for (int64_t e = 0; e < E; ++e) tp[g->out_idx[e] + 1]++;
  bool use_in = g->in_ptr && g->in_idx && g->in_ptr[0] == 0;
  for (int32_t v = 0; v < N && use_in; ++v) {
    const int64_t a = g->in_ptr[v], b = g->in_ptr[v + 1];
    use_in = b - a == tp[v + 1];
    for (int64_t e = a + 1; e < b && use_in; ++e) use_in = g->in_idx[e - 1] <= g->in_idx[e];
  }
  for (int32_t v = 0; v < N; ++v) tp[v + 1] += tp[v];
  std::vector<int32_t> built;
  const int32_t* ti = g->in_idx;
  if (!use_in) {
    built.resize((size_t)E);
    std::vector<int64_t> fill(tp.begin(), tp.end() - 1);
    for (int32_t i = 0; i < N; ++i)
      for (int64_t e = g->out_ptr[i]; e < g->out_ptr[i + 1]; ++e) built[fill[g->out_idx[e]]++] = i;
    ti = built.data();
  }
  std::vector<uint8_t> visited((size_t)N, 0);
  std::vector<int32_t> stk;
  stk.reserve((size_t)N + 16);
  int32_t b = 0;
  for (auto it = finish.rbegin(); it != finish.rend(); ++it) {
    if (visited[*it]) continue;
    int32_t count = 0;
    stk.clear();
    stk.push_back(*it);
    while (!stk.empty() && count < dfs_depth) {
      const int32_t v = stk.back();
      stk.pop_back();
      if (visited[v]) continue;
      visited[v] = 1;
      labels[v] = b;
      ++count;
      for (int64_t e = tp[v]; e < tp[v + 1]; ++e)
        if (!visited[ti[e]]) stk.push_back(ti[e]);
    }
    ++b;
  }
  *B = b;
  return USAP_OK;
}

int usap_build_csr(int32_t N, int64_t E, const int32_t* src, const int32_t* dst, const int32_t* w,
                   int64_t* out_ptr, int32_t* out_idx, int32_t* out_w, int64_t* in_ptr,
                   int32_t* in_idx, int32_t* in_w) {
  if (N < 0 || E < 0 || !out_ptr || !in_ptr || (E > 0 && (!src || !dst || !out_idx || !in_idx)))
    return USAP_EINVAL;
  for (int64_t e = 0; e < E; ++e)
    if (src[e] < 0 || src[e] >= N || dst[e] < 0 || dst[e] >= N) return USAP_EINVAL;
  std::fill(out_ptr, out_ptr + N + 1, 0);
  std::fill(in_ptr, in_ptr + N + 1, 0);
  for (int64_t e = 0; e < E; ++e) {
    out_ptr[src[e] + 1]++;
    in_ptr[dst[e] + 1]++;
  }
  for (int32_t v = 0; v < N; ++v) {
    out_ptr[v + 1] += out_ptr[v];
    in_ptr[v + 1] += in_ptr[v];
  }
  std::vector<int64_t> fo(out_ptr, out_ptr + N), fi(in_ptr, in_ptr + N);
  for (int64_t e = 0; e < E; ++e) {  // file order within each list
    const int64_t po = fo[src[e]]++, pi = fi[dst[e]]++;
    out_idx[po] = dst[e];
    in_idx[pi] = src[e];
    if (w && out_w) out_w[po] = w[e];
    if (w && in_w) in_w[pi] = w[e];
  }
  return USAP_OK;
}

// load_graph_from_tsv, graph_parallel.hpp:251-317 ("from\tto\tweight",
// 1-based, weight parsed as a number and truncated like static_cast<W>(stof)).
int usap_load_tsv(const char* path, int32_t* N, int64_t* E, int32_t* src, int32_t* dst,
                  int32_t* w) {
  if (!path || !N || !E) return USAP_EINVAL;
  FILE* f = std::fopen(path, "r");
  if (!f) return USAP_EIO;
  char line[512];
  int64_t e = 0, maxid = 0;
  const bool fill = src && dst;
  while (std::fgets(line, sizeof line, f)) {
    char* p = line;
    while (*p == ' ' || *p == '\t') ++p;
    if (*p == '\0' || *p == '\n' || *p == '#') continue;
    char* q;
    const long a = std::strtol(p, &q, 10);
    if (q == p) {
      std::fclose(f);
      return USAP_EIO;
    }
    p = q;
    const long b = std::strtol(p, &q, 10);
    if (q == p || a < 1 || b < 1 || a > 2147483647L || b > 2147483647L) {
      std::fclose(f);
      return USAP_EIO;
    }
    p = q;
    double wt = std::strtod(p, &q);
    if (q == p) wt = 1.0;
    if (fill) {
      if (e >= *E) {
        std::fclose(f);
        return USAP_EINVAL;
      }
      src[e] = (int32_t)(a - 1);
      dst[e] = (int32_t)(b - 1);
      if (w) w[e] = (int32_t)wt;
    }
    maxid = std::max<int64_t>(maxid, std::max(a, b));
    ++e;
  }
  std::fclose(f);
  if (fill && e != *E) return USAP_EINVAL;
  *E = e;
  *N = (int32_t)maxid;
  return USAP_OK;
}

int usap_generate_dcsbm(int32_t N, int32_t K, double alpha, int32_t min_block, int32_t dmin,
                        int32_t dmax, double gamma, double p_in, uint64_t seed, int64_t* E,
                        int32_t* src, int32_t* dst, int32_t* truth) {
  if (N < 2 || K < 1 || alpha <= 0 || min_block < 1 || (int64_t)K * min_block > N || dmin < 1 ||
      dmax < dmin || p_in < 0 || p_in > 1 || !E)
    return USAP_EINVAL;
  // out-degrees: P(d) ~ d^-gamma on [dmin, dmax], capped at N-1
  std::vector<double> cdf((size_t)(dmax - dmin + 1));
  double tot = 0;
  for (int32_t d = dmin; d <= dmax; ++d) cdf[d - dmin] = (tot += std::pow((double)d, -gamma));
  std::vector<int32_t> deg((size_t)N);
  Stream sd(seed, 3);
  int64_t etot = 0;
  for (int32_t u = 0; u < N; ++u) {
    const double x = sd.uniform() * tot;
    const int32_t d = dmin + (int32_t)(std::upper_bound(cdf.begin(), cdf.end(), x) - cdf.begin());
    deg[u] = std::min(std::min(d, dmax), N - 1);
    etot += deg[u];
  }
  if (!src || !dst) {
    *E = etot;
    return USAP_OK;
  }
  if (*E != etot) return USAP_EINVAL;
  // block sizes: min_block + Dirichlet(alpha) share of the rest
  Stream sb(seed, 1);
  std::vector<double> g((size_t)K);
  double gs = 0;
  for (int32_t k = 0; k < K; ++k) gs += (g[k] = sb.gamma(alpha));
  const int64_t rest = (int64_t)N - (int64_t)K * min_block;
  std::vector<int64_t> size((size_t)K);
  int64_t used = 0;
  for (int32_t k = 0; k < K; ++k) used += (size[k] = min_block + (int64_t)std::floor(g[k] / gs * rest));
  for (int32_t k = 0; used < N; k = (k + 1) % K, ++used) size[k]++;
  // node ids scattered over blocks by a random permutation
  Stream sp(seed, 2);
  std::vector<int32_t> perm((size_t)N);
  for (int32_t i = 0; i < N; ++i) perm[i] = i;
  for (int32_t i = N - 1; i > 0; --i) std::swap(perm[i], perm[sp.index(i + 1)]);
  std::vector<int32_t> lab((size_t)N);
  std::vector<std::vector<int32_t>> members((size_t)K);
  for (int32_t k = 0, j = 0; k < K; ++k)
    for (int64_t c = 0; c < size[k]; ++c, ++j) {
      lab[perm[j]] = k;
      members[k].push_back(perm[j]);
    }
  // edges, grouped by source ascending; no self-loops, no duplicates
  Stream se(seed, 4);
  std::vector<int32_t> stamp((size_t)N, -1);
  int64_t e = 0;
  for (int32_t u = 0; u < N; ++u) {
    stamp[u] = u;
    for (int32_t j = 0; j < deg[u]; ++j) {
      // target block once per edge; duplicates and self-loops redraw the node only
      int32_t tb = lab[u];
      if (K > 1 && se.uniform() >= p_in) tb = (int32_t)((lab[u] + 1 + se.index(K - 1)) % K);
      const auto& m = members[tb];
      int32_t v = -1;
      for (int attempt = 0; attempt < 64 && v < 0; ++attempt) {
        const int32_t c = m[se.index((int64_t)m.size())];
        if (stamp[c] != u) v = c;
      }
      while (v < 0) {  // block exhausted: any node
        const int32_t c = (int32_t)se.index(N);
        if (stamp[c] != u) v = c;
      }
      stamp[v] = u;
      src[e] = u;
      dst[e] = v;
      ++e;
    }
  }
  if (truth) std::memcpy(truth, lab.data(), sizeof(int32_t) * (size_t)N);
  return USAP_OK;
}

}  // extern "C"
