// engine.hip -- MI355X (gfx950) engine for uSAP's partition hot path.
//
// Device state per graph: CSR/CSC (int64 offsets, int32 ids), labels int32[N],
// block degrees int32[B], and the interblock edge-count matrix M as one "block
// row" per block a holding both orientations: the slot of key b carries
// o = M[a][b] (edges a -> b) and i = M[b][a] (edges b -> a), so every M[a][b]
// is stored twice (as o in row a and as i in row b) and one probe of row a
// answers both the row and the column lookup of a block pair.  Rows live in
// one slab pool:
//   row a occupies slots [off[a], off[a]+cap[a]);
//   cap[a] == B  -> direct-mapped row (slot = key), used when a hashed row
//                   would need >= B slots (dense regime, small B);
//   cap[a] <  B  -> open-addressing hash (power of two, linear probing);
//                   a slot is 16 bytes (key, -, o, i; key -1 = empty), so a
//                   probe and the count update after it share a cache line
//                   and one 64-bit atomic updates both counts.
// Keys are never deleted inside a sweep (counts may drop to 0).  Each row
// also keeps kNBK bucket sums (sum of o + i per key range) for the two-level
// inverse-CDF pick over the combined weights M[a][b] + M[b][a].  An
// increment that finds no free slot near its key during nodal batches is
// parked in an overflow buffer; between two batches k_grow finds the rows
// whose key count passed 1/slack of the slab or that have parked increments,
// compacts them in place or moves them to a larger slab at the pool's tail,
// and replays the parked increments.
// The host rebuilds M from the labels only when the pool runs out.
//
// Kernels (one wavefront of 64 lanes per proposal item):
//   k_degrees / k_caps / k_fill / k_insert_* / k_build_rows : M build from
//        labels (graph_parallel.hpp:703-759)
//   k_merge   : K merge proposals per block + best per block
//        (graph_parallel.hpp:349-368, :761-988), one workgroup per block,
//        one wavefront per proposal
//   k_nodal   : one nodal proposal per node against the frozen M
//        (graph_parallel.hpp:990-1222), one wavefront per node
//   k_apply   : accepted moves -> M, degrees (atomics) and labels; equals
//        the reference's full rebuild (graph_parallel.hpp:438-443); its
//        block 0 runs the sweep's update step and convergence test
//        (graph_parallel.hpp:433-458) from k_nodal's per-workgroup shares,
//        so a whole sweep runs without the host
//   k_grow    : row growth / compaction and overflow replay between batches;
//        acts on the batch's error and rebuild flags
//   k_entropy : data_S over all non-zeros (graph_parallel.hpp:1281-1319)
//
// The nodal and apply kernels are latency bound (a few hundred bytes per node
// at random rows of M); each wavefront keeps all loads of one step in flight
// at once -- list ids, then labels, then rows, then the first probes of every
// lookup -- instead of paying one memory latency per lookup.
//
// dS uses the factorised form of the reference's term sums: terms whose count
// and degrees do not change cancel exactly in integer arithmetic, so a nodal
// proposal costs O(distinct neighbour blocks) lookups and a merge proposal
// O(min row length) lookups, instead of the reference's O(B) scans.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <rocprim/rocprim.hpp>
#include <string>
#include <vector>

#include "engine_internal.hpp"
#include "usap.h"
#include "usap_common.hpp"

using usap::i128;
using usap::u128;

namespace {

constexpr int kWave = 64;
constexpr int kNodalWaves = 4;           // wavefronts per workgroup in k_nodal / k_apply
constexpr int kTab = 512;                // per-wavefront LDS table slots
constexpr int kTabMaxDeg = kTab * 3 / 4; // larger nodes use a global-memory table
constexpr int kMergeWaves = 4;           // wavefronts per workgroup in k_merge
constexpr int kNBK = 512;                // bucket sums per block row
constexpr int kLookupRatio = 8;          // a hashed slab this many times a bucket is looked up per key
constexpr int kScr = 3 * kTab;           // per-wavefront LDS scratch (ints) for one bucket
constexpr int kDegLds = 1024;            // k_degrees privatises its counters for B <= this
constexpr int kSmallB = 128;             // k_insert_small: B*B counts in LDS
constexpr int kEntropyBlocks = 4096;  // k_entropy grid at most
// Partials of the allocation-free MDL (usap_entropy_noalloc): 2 words per
// k_entropy workgroup, 64 KiB per GPU.
static __device__ __align__(256) long scratch[2 * kEntropyBlocks];
constexpr int kNumSnap = 4;

enum : int32_t {
  kFlagLoad = 1,      // a hashed row passed 3/4 load
  kFlagFull = 2,      // an insertion found no free slot (increment lost)
  kFlagMissing = 4,   // a decrement did not find its key (bug)
  kFlagNegative = 8,  // a negative count reached an entropy term (bug)
};

struct Status {
  int32_t flags;  // kFlag* raised by the running batch
  int32_t pad[3];
};

// Device state of one nodal sweep (the reference's `update` task,
// graph_parallel.hpp:433-458, run by block 0 of each batch's k_apply).
struct Sweep {
  unsigned long long cum_lo, cum_hi;  // exact cumulative dS of the sweep (int128)
  unsigned long long stream0;         // random stream of the sweep's first batch (pre-drawn: node 0's draw)
  double S;                           // running S: S += dS_sum after every batch
  double mean;                        // moving average of earlier sweeps' dS
  double thr;                         // delta_entropy_threshold1 or 2
  double dS_last;                     // dS_sum after the last batch (itr_delta_entropy[itr])
  unsigned long long moves;
  unsigned long long nodal_bytes, apply_bytes;
  int32_t cum_inf;  // an accepted dS was -inf
  int32_t check;    // itr >= window - 1: the convergence test is active
  int32_t stop;     // 0 run, 1 converged, 2 M needs a rebuild, 3 inconsistent M
  int32_t done;     // batches executed
  int32_t flags;    // OR of the batches' flags
  int32_t stop_n;   // stop as the running batch's k_nodal saw it (0: the batch runs)
};

struct GraphView {
  int32_t N;
  const int64_t* out_ptr;
  const int32_t* out_idx;
  const int32_t* out_w;
  const int64_t* in_ptr;
  const int32_t* in_idx;
  const int32_t* in_w;
};

// One slot of a block row slab: key (-1 empty; unused in a direct row) and
// the pair's two counts, o = M[a][key] and i = M[key][a].  (o, i) is 8-byte
// aligned: one 64-bit atomic adds to both.
struct __align__(16) Slot {
  int32_t k, pad;
  int32_t o, i;
};

// The two counts of a block pair (a, b): o = M[a][b], i = M[b][a].
struct Cnt {
  int32_t o, i;
};

// M as block rows, as seen by kernels.
struct MView {
  const int64_t* off;
  const int32_t* cap;
  int32_t* nkeys;
  Slot* kv;        // row slabs
  int32_t* bs;     // [B][kNBK]: sum of the row's o + i per key bucket
  int32_t* gpend;  // [B]: increments of the row waiting in the overflow buffer
  int32_t* ovf;    // [4][ovf_cap]: (row, key, do, di) increments that found no free slot
  int32_t ovf_cap;
  struct GrowState* gst;
  int32_t B;
  int32_t bshift;  // bucket of key k = k >> bshift
  int32_t slack;   // slab_slack: a hashed row grows past 1/slack load
};

// Rows to grow between two batches; slabs for grown rows come from the free
// tail of the pool.
struct GrowState {
  int32_t n;                  // unused (rows to grow are found by k_grow's scan)
  int32_t nospace;            // the pool ran out: the host rebuilds M
  int32_t ovf_n;              // entries in the overflow buffer
  unsigned long long used;    // pool slots in use
  unsigned long long grows;   // rows grown or compacted since the engine was created
  unsigned long long full;    // insertions that found no free slot
  unsigned long long nospace_events;
  unsigned long long compactions;
};

__device__ __forceinline__ uint32_t hash32(int32_t b) {
  uint32_t x = (uint32_t)b * 0x9E3779B1u;
  return x ^ (x >> 15);
}

__device__ __forceinline__ int32_t ld_relaxed(int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Slab of one row: offset and capacity.
struct Row {
  int64_t o;
  int32_t c;
};

__device__ __forceinline__ Row m_row(const MView& m, int32_t a) { return Row{m.off[a], m.cap[a]}; }

// First slot of key b in row R (the column itself in a direct row).
__device__ __forceinline__ int64_t m_slot(const MView& m, Row R, int32_t b) {
  return R.c == m.B ? R.o + b : R.o + (int64_t)(hash32(b) & (uint32_t)(R.c - 1));
}

// A lookup's first probe: the slot with the key and count found there.  The
// callers issue several probes before resolving any, so a wavefront has all
// of its lookups in flight at once instead of one memory latency per lookup.
struct Probe {
  uint32_t h;  // slot within the row
  int32_t k, o, i;
};

__device__ __forceinline__ Probe m_probe(const MView& m, Row R, int32_t b) {
  Probe p;
  p.h = R.c == m.B ? (uint32_t)b : hash32(b) & (uint32_t)(R.c - 1);
  const Slot q = m.kv[R.o + p.h];
  p.k = q.k;
  p.o = q.o;
  p.i = q.i;
  return p;
}

// Linear probing on from slot `slot` of a hashed row after a collision there.
__device__ __forceinline__ Cnt m_value_more(const Slot* kv, int64_t o, int32_t c, int32_t b,
                                            int64_t slot) {
  const uint32_t mask = (uint32_t)c - 1;
  uint32_t h = (uint32_t)(slot - o);
  for (int32_t probe = 1; probe < c; ++probe) {
    h = (h + 1) & mask;
    const Slot q = kv[o + h];
    if (q.k == b) return Cnt{q.o, q.i};
    if (q.k < 0) return Cnt{0, 0};
  }
  return Cnt{0, 0};
}

// (M[a][b], M[b][a]) from the first probe p of key b in row R of a.
__device__ __forceinline__ Cnt m_value(const MView& m, Row R, int32_t b, const Probe& p) {
  if (R.c == m.B || p.k == b) return Cnt{p.o, p.i};
  if (p.k < 0) return Cnt{0, 0};
  return m_value_more(m.kv, R.o, R.c, b, R.o + p.h);
}

// (M[a][b], M[b][a])
__device__ __forceinline__ Cnt m_get(const MView& m, int32_t a, int32_t b) {
  const Row R = m_row(m, a);
  return m_value(m, R, b, m_probe(m, R, b));
}

// o += dout, i += din with one 64-bit atomic.  Exact as long as neither count
// leaves [0, 2^31) at any point, which holds for every caller: a decrement
// only removes edges the count held when the kernel started.
__device__ __forceinline__ void cnt_add(Slot* q, int32_t dout, int32_t din) {
  const long long d = (long long)din * 4294967296LL + (long long)dout;
  atomicAdd(reinterpret_cast<unsigned long long*>(&q->o), (unsigned long long)d);
}

constexpr int32_t kMaxWalk = 16;  // probes before a nodal-batch increment is parked

// M[a][b] += dout and M[b][a] += din in row R of a (negative only for a key
// that exists), given the key k read at the first probed slot `slot`.  Keys
// only go from empty to set, so an earlier read is a valid first read: a
// stale "empty" makes the claim fail and return the key that is there.
// kBuckets also updates the row's bucket sum, and leaves a crowded row to
// k_grow (by its key count, or a parked increment) instead of flagging a
// rebuild.
template <bool kBuckets>
__device__ __forceinline__ void m_add_at(const MView& m, int32_t a, Row R, int32_t b, int32_t dout,
                                         int32_t din, int64_t slot, int32_t k, int32_t* flags) {
  if (kBuckets && dout + din != 0)
    atomicAdd(&m.bs[(int64_t)a * kNBK + (b >> m.bshift)], dout + din);
  if (R.c == m.B || k == b) {
    cnt_add(&m.kv[slot], dout, din);
    return;
  }
  const bool dec = dout < 0 || din < 0;
  const int64_t o = R.o;
  const int32_t c = R.c;
  const uint32_t mask = (uint32_t)c - 1;
  uint32_t h = (uint32_t)(slot - o);
  for (int32_t probe = 0; probe < c; ++probe) {
    // an increment whose key is still absent after kMaxWalk slots is parked
    // (k_grow grows the row) rather than walking a crowded row
    if (kBuckets && !dec && probe >= kMaxWalk) break;
    if (probe > 0) k = ld_relaxed(&m.kv[o + h].k);
    if (k < 0) {
      if (dec) {
        atomicOr(flags, kFlagMissing);
        return;
      }
      k = atomicCAS(&m.kv[o + h].k, -1, b);
      if (k == -1) {
        if (kBuckets) {
          // the counts and the key count without waiting for either: k_grow
          // finds the rows past 1/slack load (short probe chains for the
          // lookups, most of which miss) from the key counts themselves
          cnt_add(&m.kv[o + h], dout, din);
          atomicAdd(&m.nkeys[a], 1);
          return;
        }
        // past 3/4 load in a build: rebuild
        if (4 * (atomicAdd(&m.nkeys[a], 1) + 1) > 3 * c) atomicOr(flags, kFlagLoad);
        k = b;
      }
    }
    if (k == b) {
      cnt_add(&m.kv[o + h], dout, din);
      return;
    }
    h = (h + 1) & mask;
  }
  if (kBuckets && !dec) {  // no free slot near: park the increment until k_grow has grown the row
    atomicAdd(&m.gst->full, 1ull);
    const int32_t i = atomicAdd(&m.gst->ovf_n, 1);
    if (i < m.ovf_cap) {
      m.ovf[i] = a;
      m.ovf[m.ovf_cap + i] = b;
      m.ovf[2 * m.ovf_cap + i] = dout;
      m.ovf[3 * m.ovf_cap + i] = din;
      atomicAdd(&m.gpend[a], 1);  // k_grow grows every row with parked increments
      return;
    }
  }
  atomicOr(flags, dec ? kFlagMissing : kFlagFull);
}

// Where m_add_at starts for key b in row R: the keys of kWin slots from b's
// home slot are read at once (one memory latency for most probe chains) and
// the first that holds b or is empty is returned with its slot; if none does,
// the window's last slot, from which m_add_at walks on.
constexpr int kWin = 8;
__device__ __forceinline__ int32_t m_first_key(const MView& m, Row R, int32_t b, int64_t& slot) {
  if (R.c == m.B) {
    slot = R.o + b;
    return b;
  }
  const uint32_t mask = (uint32_t)R.c - 1;  // hashed rows hold >= 32 slots
  const uint32_t h = hash32(b) & mask;
  int32_t k[kWin];
#pragma unroll
  for (int q = 0; q < kWin; ++q) k[q] = ld_relaxed(&m.kv[R.o + ((h + q) & mask)].k);
  int32_t kk = k[kWin - 1];
  uint32_t pos = (h + kWin - 1) & mask;
  bool found = false;
#pragma unroll
  for (int q = 0; q < kWin; ++q) {
    if (!found && (k[q] == b || k[q] < 0)) {
      found = true;
      kk = k[q];
      pos = (h + q) & mask;
    }
  }
  slot = R.o + pos;
  return kk;
}

// M[a][b] += dout, M[b][a] += din in row a
template <bool kBuckets>
__device__ __forceinline__ void m_add(const MView& m, int32_t a, int32_t b, int32_t dout,
                                      int32_t din, int32_t* flags) {
  const Row R = m_row(m, a);
  int64_t slot;
  const int32_t k = m_first_key(m, R, b, slot);
  m_add_at<kBuckets>(m, a, R, b, dout, din, slot, k, flags);
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// Draw d of a pre-drawn uniform stream: its four words, slot order as the
// Philox outputs they replace.
__device__ __forceinline__ void predrawn(const uint64_t* uni, uint64_t d, uint64_t x[4]) {
  const ulonglong2* q = reinterpret_cast<const ulonglong2*>(uni + 4 * d);
  const ulonglong2 a = q[0], b = q[1];
  x[0] = a.x;
  x[1] = a.y;
  x[2] = b.x;
  x[3] = b.y;
}

// Orders this wavefront's earlier table writes before its later reads.
__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ long wave_sum(long v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// v of lane l (l the same in every lane)
__device__ __forceinline__ long read_lane(long v, int l) {
  const unsigned long long u = (unsigned long long)v;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
  return (long)(((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ i128 wave_sum128(i128 v) {
  unsigned long lo = (unsigned long)(u128)v;
  long hi = (long)(v >> 64);
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const unsigned long lo2 = __shfl_xor(lo, off);
    const long hi2 = __shfl_xor(hi, off);
    const unsigned long s = lo + lo2;
    hi = hi + hi2 + (s < lo ? 1 : 0);
    lo = s;
  }
  return (i128)(((u128)(unsigned long)hi << 64) | (u128)lo);
}

__device__ __forceinline__ long wave_incl_scan(long v) {
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const long t = __shfl_up(v, off);
    if (lane >= off) v += t;
  }
  return v;
}

// The terms read the lambda table unconditionally (lambda of m <= 0 reads
// lam[0] = 0) and select the result, so the table loads of a whole dS are
// issued together: a load under `if (m > 0)` would sit in its own branch and
// wait for every earlier term's loads first.
__device__ __forceinline__ int64_t lam_of(const int64_t* lam, int64_t m) {
  return lam[m > 0 ? m : 0];
}

// m * lambda(m): one L term
__device__ __forceinline__ i128 Lterm(const int64_t* lam, int64_t m) {
  const int64_t L = lam_of(lam, m);
  return m > 0 ? (i128)m * (i128)L : (i128)0;
}

// m * lambda(d): one degree-factor term; inf when m > 0 and d == 0
__device__ __forceinline__ i128 Pterm(const int64_t* lam, int64_t m, int64_t d, bool& inf) {
  const int64_t L = lam_of(lam, d);
  inf = inf || (m > 0 && d <= 0);
  return (m > 0 && d > 0) ? (i128)m * (i128)L : (i128)0;
}

// T(m, x, y) = m*(lambda(m)-lambda(x)-lambda(y)); inf when m > 0 and x*y == 0
__device__ __forceinline__ i128 Tterm(const int64_t* lam, int64_t m, int64_t x, int64_t y,
                                      bool& inf) {
  const int64_t Lm = lam_of(lam, m), Lx = lam_of(lam, x), Ly = lam_of(lam, y);
  inf = inf || (m > 0 && (x <= 0 || y <= 0));
  return (m > 0 && x > 0 && y > 0) ? (i128)m * (i128)(Lm - Lx - Ly) : (i128)0;
}

// per-lane part of sum_{lo <= key <= x, key != excl} M[a][key] + M[key][a]
// over row (o, c) of a
__device__ __forceinline__ long count_range_lane(const MView& m, int64_t o, int32_t c, int32_t lo,
                                                 int32_t x, int32_t excl) {
  const int lane = lane_id();
  long acc = 0;
  if (c == m.B) {
    for (int32_t i = lo + lane; i <= x; i += kWave)
      if (i != excl) acc += (long)m.kv[o + i].o + m.kv[o + i].i;
  } else {
    for (int32_t j = lane; j < c; j += kWave) {
      const Slot q = m.kv[o + j];
      if (q.k >= lo && q.k <= x && q.k != excl) acc += (long)q.o + q.i;
    }
  }
  return acc;
}

// scr[k - lo] += M[a][k] + M[k][a] for lo <= k < hi, from row R of a: a
// direct row is read at keys [lo, hi), a hashed row's slab is scanned eight
// slots per lane at a time (all loads of a pass in flight before the LDS adds).
__device__ __forceinline__ void gather_row(const MView& m, Row R, int32_t lo, int32_t hi,
                                           int32_t* scr, long& bytes) {
  const int lane = lane_id();
  if (R.c == m.B) {
    for (int32_t i = lane; i < hi - lo; i += kWave) {
      const Slot q = m.kv[R.o + lo + i];
      atomicAdd(&scr[i], q.o + q.i);
    }
    bytes += 8 * (long)(hi - lo);
    return;
  }
  for (int32_t base = 0; base < R.c; base += 8 * kWave) {
    int32_t k[8], v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int32_t j = base + q * kWave + lane;
      k[q] = -1;
      v[q] = 0;
      if (j < R.c) {
        const Slot x = m.kv[R.o + j];
        k[q] = x.k;
        v[q] = x.o + x.i;
      }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (k[q] >= lo && k[q] < hi && v[q] != 0) atomicAdd(&scr[k[q] - lo], v[q]);
  }
  bytes += 12 * (long)R.c;
}

// scr[k - lo] = M[a][k] + M[k][a] for lo <= k < hi from hashed row R of a, by
// looking each key up: four keys per lane at a time, every pending walk of
// the wavefront advancing one slot per step with all its loads in flight.  A
// walk ends at its key or at an empty slot, and after R.c slots at the latest
// (a full slab).  For a slab much larger than the bucket this reads a few
// slots per key instead of the whole slab.
__device__ __forceinline__ void lookup_bucket(const MView& m, Row R, int32_t lo, int32_t hi,
                                              int32_t* scr, long& bytes) {
  const int lane = lane_id();
  const uint32_t mask = (uint32_t)R.c - 1;
  for (int32_t base = lo; base < hi; base += 4 * kWave) {
    uint32_t h[4];
    int32_t wv[4];
    uint32_t pend = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int32_t key = base + q * kWave + lane;
      h[q] = hash32(key) & mask;
      wv[q] = 0;
      if (key < hi) pend |= 1u << q;
    }
    for (int32_t step = 0; step < R.c && __ballot(pend != 0); ++step) {
      Slot x[4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if ((pend >> q) & 1) x[q] = m.kv[R.o + h[q]];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        bytes += 16 * (long)__popcll(__ballot((pend >> q) & 1));
        if (!((pend >> q) & 1)) continue;
        const int32_t key = base + q * kWave + lane;
        if (x[q].k == key) {
          wv[q] = x[q].o + x[q].i;
          pend &= ~(1u << q);
        } else if (x[q].k < 0) {
          pend &= ~(1u << q);
        } else {
          h[q] = (h[q] + 1) & mask;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int32_t key = base + q * kWave + lane;
      if (key < hi) scr[key - lo] = wv[q];
    }
  }
}

// Index of the t-th unit in the LDS weights scr[0, width) (t < their sum):
// each lane sums a contiguous run of entries, one wavefront scan finds the
// run, and the lane owning it walks it.
__device__ __forceinline__ int32_t scan_scratch(const int32_t* scr, int32_t width, int64_t t) {
  const int lane = lane_id();
  const int32_t per = (width + kWave - 1) / kWave;
  const int32_t b0 = lane * per;
  long sum = 0;
  for (int32_t k = 0; k < per; ++k)
    if (b0 + k < width) sum += scr[b0 + k];
  const long inc = wave_incl_scan(sum);
  const int f = __ffsll((unsigned long long)__ballot(inc > t)) - 1;
  if (f < 0) return width - 1;
  long tl = t - (__shfl(inc, f) - __shfl(sum, f));
  int32_t res = width - 1;
  if (lane == f) {
    for (int32_t k = 0; k < per && b0 + k < width; ++k) {
      const long w = scr[b0 + k];
      if (tl < w) {
        res = b0 + k;
        break;
      }
      tl -= w;
    }
  }
  return __shfl(res, f);
}

// Inverse-CDF pick over the combined weights M[u][i] + M[i][u] (block row
// u: o + i), ascending i, skipping i == excl: the first i whose running sum
// exceeds t (graph_parallel.hpp:849-850 / :1061-1062 with exact integer
// weights).  The caller guarantees t < total weight.  Two levels: the row's
// bucket sums (all loaded at once) locate the key bucket [lo, hi); the
// bucket's weights then come from one load per key when the row is direct and
// the bucket is at most a wavefront wide, else they are gathered into the
// wavefront's LDS scratch `scr` (scr_cap ints) -- a hashed row by scanning its
// slab, which costs a few loads per lane where looking up each absent key
// would walk a probe chain, or, when the slab holds kLookupRatio times the
// bucket's keys or more, by looking each key up (a whole-slab scan of a
// 16384-slot row took one wavefront ~120 us) -- and searched with one scan.
__device__ __forceinline__ int32_t select_combined(const MView& M, int32_t u, int64_t t,
                                                   int32_t excl, int32_t* scr, int32_t scr_cap,
                                                   long& bytes) {
  const int lane = lane_id();
  const int32_t B = M.B;
  const int sh = M.bshift;
  const Row RM = m_row(M, u);
  const int32_t nbk = ((B - 1) >> sh) + 1;
  const int32_t* bM = M.bs + (int64_t)u * kNBK;
  constexpr int kQ = kNBK / kWave;
  long wq[kQ];
#pragma unroll
  for (int q = 0; q < kQ; ++q) {
    const int32_t j = q * kWave + lane;
    wq[q] = j < nbk ? (long)bM[j] : 0;
  }
  bytes += 12 + 4 * (long)nbk;
  if (excl >= 0) {
    const Probe p = m_probe(M, RM, excl);
    const Cnt x = m_value(M, RM, excl, p);
    const long wx = (long)x.o + x.i;
    bytes += 12;
    const int32_t xb = excl >> sh;
#pragma unroll
    for (int q = 0; q < kQ; ++q)
      if (q * kWave + lane == xb) wq[q] -= wx;
  }
  // level 1: bucket jb holding the t-th unit
  int32_t jb = nbk - 1;
#pragma unroll
  for (int q = 0; q < kQ; ++q) {
    const long inc = wave_incl_scan(wq[q]);
    const long tot = __shfl(inc, kWave - 1);
    if (t < tot) {
      const int f = __ffsll((unsigned long long)__ballot(inc > t)) - 1;
      jb = q * kWave + f;
      t -= __shfl(inc, f) - __shfl(wq[q], f);
      break;
    }
    t -= tot;
  }
  // level 2: keys [lo, hi) of that bucket in ascending order
  const int32_t lo = jb << sh;
  const int32_t hi = min(B, lo + (1 << sh));
  const int32_t width = hi - lo;
  // a bucket of one key (every bucket when B <= kNBK): the bucket's sum is
  // that key's weight, so the bucket is the pick (the scans below would
  // return lo too, also for the excluded key's emptied bucket)
  if (width == 1) return lo;
  if (width <= kWave && RM.c == B) {  // direct row: one load per key
    long w = 0;
    const int32_t i = lo + lane;
    if (lane < width && i != excl) {
      const Slot q = M.kv[RM.o + i];
      w = (long)q.o + q.i;
    }
    bytes += 8 * (long)width;
    const long inc = wave_incl_scan(w);
    const int f = __ffsll((unsigned long long)__ballot(inc > t)) - 1;
    return f < 0 ? hi - 1 : lo + f;
  }
  if (width <= scr_cap) {
    if (RM.c == B) {
      for (int32_t base = 0; base < width; base += 4 * kWave) {
        int32_t v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int32_t i = base + q * kWave + lane;
          v[q] = 0;
          if (i < width) {
            const Slot x = M.kv[RM.o + lo + i];
            v[q] = x.o + x.i;
          }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int32_t i = base + q * kWave + lane;
          if (i < width) scr[i] = v[q];
        }
      }
      bytes += 8 * (long)width;
    } else if (RM.c >= kLookupRatio * width) {
      lookup_bucket(M, RM, lo, hi, scr, bytes);
    } else {
      for (int32_t i = lane; i < width; i += kWave) scr[i] = 0;
      wave_fence();
      gather_row(M, RM, lo, hi, scr, bytes);
    }
    wave_fence();
    if (lane == 0 && excl >= lo && excl < hi) scr[excl - lo] = 0;
    wave_fence();
    return lo + scan_scratch(scr, width, t);
  }
  // buckets wider than the scratch: bisection inside the bucket
  int32_t a = lo, b = hi - 1;
  while (a < b) {
    const int32_t mid = a + (b - a) / 2;
    const long c = wave_sum(count_range_lane(M, RM.o, RM.c, lo, mid, excl));
    if (c > t)
      b = mid;
    else
      a = mid + 1;
  }
  return a;
}

// t-th unit of the node's neighbour list (out-neighbours then in-neighbours,
// weight w), graph_parallel.hpp:1010-1027.
__device__ int32_t pick_neighbor(const GraphView& g, int64_t o0, int64_t kco, int64_t i0,
                                 int64_t kci, int64_t t) {
  if (!g.out_w) return t < kco ? g.out_idx[o0 + t] : g.in_idx[i0 + (t - kco)];
  const int lane = lane_id();
  const int64_t kc = kco + kci;
  long acc = 0;
  for (int64_t base = 0; base < kc; base += kWave) {
    const int64_t j = base + lane;
    long w = 0;
    if (j < kco)
      w = g.out_w[o0 + j];
    else if (j < kc)
      w = g.in_w[i0 + (j - kco)];
    const long inc = wave_incl_scan(w);
    const long tot = __shfl(inc, kWave - 1);
    if (acc + tot > t) {
      const unsigned long long hit = __ballot(acc + inc > t);
      const int64_t jj = base + (__ffsll((unsigned long long)hit) - 1);
      return jj < kco ? g.out_idx[o0 + jj] : g.in_idx[i0 + (jj - kco)];
    }
    acc += tot;
  }
  return -1;
}

// ---------------------------------------------------------------------------
// M build
// ---------------------------------------------------------------------------

// Block degrees, edge counts and node counts from the labels; counters are
// privatised in LDS when B is small (many nodes share a block).
__global__ __launch_bounds__(256) void k_degrees(GraphView g, const int32_t* __restrict__ labels,
                                                 int32_t B, int32_t* d_out, int32_t* d_in,
                                                 int32_t* ecnt_out, int32_t* ecnt_in,
                                                 int32_t* bsize) {
  __shared__ int32_t s_c[5][kDegLds];
  const bool lds = B <= kDegLds;
  if (lds) {
    for (int i = threadIdx.x; i < 5 * kDegLds; i += blockDim.x) (&s_c[0][0])[i] = 0;
    __syncthreads();
  }
  int32_t* dst[5] = {d_out, d_in, ecnt_out, ecnt_in, bsize};
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < g.N;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int32_t a = labels[x];
    const int64_t o0 = g.out_ptr[x], o1 = g.out_ptr[x + 1];
    const int64_t i0 = g.in_ptr[x], i1 = g.in_ptr[x + 1];
    int64_t wo = o1 - o0, wi = i1 - i0;
    if (g.out_w) {
      wo = 0;
      wi = 0;
      for (int64_t e = o0; e < o1; ++e) wo += g.out_w[e];
      for (int64_t e = i0; e < i1; ++e) wi += g.in_w[e];
    }
    const int32_t v[5] = {(int32_t)wo, (int32_t)wi, (int32_t)(o1 - o0), (int32_t)(i1 - i0), 1};
    for (int k = 0; k < 5; ++k) {
      if (!v[k]) continue;
      if (lds)
        atomicAdd(&s_c[k][a], v[k]);
      else
        atomicAdd(&dst[k][a], v[k]);
    }
  }
  if (lds) {
    __syncthreads();
    for (int32_t b = threadIdx.x; b < B; b += blockDim.x)
      for (int k = 0; k < 5; ++k)
        if (s_c[k][b]) atomicAdd(&dst[k][b], s_c[k][b]);
  }
}

// Slab capacity of a row with ecnt edges: a power of two >= slack*min(B, ecnt)+16
// (so keys inserted by later moves fit without a rebuild), or a direct row of
// B slots once that would reach B/8 (bounds a hashed row's scan; direct rows
// total at most 8*slack*E slots).
__device__ __forceinline__ int32_t cap_for(int32_t B, int32_t ecnt, int32_t slack) {
  const int64_t bound = ecnt < B ? ecnt : B;
  int64_t c = 32;
  while (c < slack * bound + 16) c <<= 1;
  return 8 * c >= B ? B : (int32_t)c;
}

// Block row a holds the keys of its out-edges and of its in-edges: its slab
// is sized for ecnt_out[a] + ecnt_in[a] keys.
__global__ void k_caps(int32_t B, int32_t slack, const int32_t* ecnt_out, const int32_t* ecnt_in,
                       int32_t* cap, int64_t* tmp) {
  const int32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a > B) return;
  if (a == B) {
    tmp[B] = 0;
    return;
  }
  const int64_t ec = (int64_t)ecnt_out[a] + ecnt_in[a];
  const int32_t c = cap_for(B, (int32_t)min<int64_t>(ec, (int64_t)B), slack);
  cap[a] = c;
  tmp[a] = c;
}

__global__ void k_fill(Slot* kv, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    kv[i] = Slot{-1, 0, 0, 0};
  }
}

// Edge u -> v between blocks a -> b: o of row a at key b and i of row b at
// key a (one update when a == b).
// A row that k_build_rows lays out in LDS: hashed, at most kRowTab slots.
constexpr int kRowTab = 1024;
__device__ __forceinline__ bool lds_row(int32_t cap, int32_t B) { return cap < B && cap <= kRowTab; }

// kSkipLds: only the rows k_build_rows does not build (direct, or hashed
// above kRowTab slots), which may hold many nodes: edge by edge keeps them
// load-balanced.
template <bool kSkipLds>
__global__ void k_insert_edges(int64_t E, const int32_t* __restrict__ src,
                               const int32_t* __restrict__ dst, const int32_t* __restrict__ w,
                               const int32_t* __restrict__ labels, MView M, int32_t* flags) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t a = labels[src[e]];
    const int32_t b = labels[dst[e]];
    const int32_t ww = w ? w[e] : 1;
    const bool ia = !kSkipLds || !lds_row(M.cap[a], M.B);
    const bool ib = !kSkipLds || !lds_row(M.cap[b], M.B);
    if (a == b) {
      if (ia) m_add<false>(M, a, a, ww, ww, flags);
    } else {
      if (ia) m_add<false>(M, a, b, ww, 0, flags);
      if (ib) m_add<false>(M, b, a, 0, ww, flags);
    }
  }
}

// Nodes grouped by block (for the per-row build): node x goes to
// nodes[off[labels[x]] + k], k its arrival order in its block.
__global__ void k_node_lists(int32_t N, const int32_t* __restrict__ labels,
                             const int32_t* __restrict__ off, int32_t* cur, int32_t* nodes) {
  for (int32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < N; x += gridDim.x * blockDim.x) {
    const int32_t a = labels[x];
    nodes[off[a] + atomicAdd(&cur[a], 1)] = x;
  }
}

// M build for large B: each hashed row of at most kRowTab slots by one
// wavefront, from the lists of the row's own nodes alone -- an out-edge
// x -> v adds to M[a][label v] (o), an in-edge u -> x to M[label u][a] (i) --
// laid out in LDS and written out whole (a slab that size holds at most
// (kRowTab - 16) / 2 list entries, so no wavefront has a long row).  The
// per-edge build sends every edge to two rows' slabs at random instead; it
// still builds the other rows (k_insert_edges<true>).
__global__ __launch_bounds__(kWave* kNodalWaves) void k_build_rows(
    GraphView g, const int32_t* __restrict__ labels, const int32_t* __restrict__ bl_off,
    const int32_t* __restrict__ bl_nodes, MView M, int32_t* flags) {
  __shared__ int32_t s_k[kNodalWaves][kRowTab];
  __shared__ int32_t s_o[kNodalWaves][kRowTab];
  __shared__ int32_t s_i[kNodalWaves][kRowTab];
  __shared__ int32_t s_pre[kNodalWaves][kWave + 1];
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  const int32_t a = blockIdx.x * kNodalWaves + w;
  if (a >= M.B) return;  // no workgroup barrier below
  const int64_t o = M.off[a];
  const int32_t c = M.cap[a];
  if (!lds_row(c, M.B)) return;
  int32_t* tk = s_k[w];
  int32_t* to = s_o[w];
  int32_t* ti = s_i[w];
  int32_t* pre = s_pre[w];
  for (int32_t j = lane; j < c; j += kWave) {
    tk[j] = -1;
    to[j] = 0;
    ti[j] = 0;
  }
  const uint32_t mask = (uint32_t)c - 1;
  const int32_t n0 = bl_off[a], n1 = a + 1 < M.B ? bl_off[a + 1] : g.N;
  long nk = 0;  // keys this lane placed in the LDS slab
  bool full = false;
  for (int32_t base = n0; base < n1; base += kWave) {  // 64 nodes at a time
    const int32_t j = base + lane;
    int64_t o0 = 0, i0 = 0;
    long ko = 0, kn = 0;
    if (j < n1) {
      const int32_t x = bl_nodes[j];
      o0 = g.out_ptr[x];
      i0 = g.in_ptr[x];
      ko = g.out_ptr[x + 1] - o0;
      kn = ko + (g.in_ptr[x + 1] - i0);
    }
    const long inc = wave_incl_scan(kn);
    const long tot = __shfl(inc, kWave - 1);
    wave_fence();  // the previous chunk's reads of pre are done
    pre[lane + 1] = (int32_t)inc;
    if (lane == 0) pre[0] = 0;
    wave_fence();
    for (long e0 = 0; e0 < tot; e0 += kWave) {
      const int32_t ei = (int32_t)(e0 + lane);
      int lo = 0, hi = kWave - 1;  // the node k of entry ei: pre[k] <= ei < pre[k + 1]
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= ei)
          lo = mid;
        else
          hi = mid - 1;
      }
      const int64_t xo = __shfl(o0, lo), xi = __shfl(i0, lo);
      const long xko = __shfl(ko, lo);
      if (ei >= tot) continue;
      const long r = ei - pre[lo];
      const bool out = r < xko;
      int32_t v, wt = 1;
      if (out) {
        v = g.out_idx[xo + r];
        if (g.out_w) wt = g.out_w[xo + r];
      } else {
        v = g.in_idx[xi + (r - xko)];
        if (g.in_w) wt = g.in_w[xi + (r - xko)];
      }
      const int32_t b = labels[v];
      uint32_t h = hash32(b) & mask;
      bool done = false;
      for (int32_t probe = 0; probe < c && !done; ++probe) {
        int32_t k = __hip_atomic_load(&tk[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        if (k < 0) {
          k = atomicCAS(&tk[h], -1, b);
          if (k == -1) {
            k = b;
            ++nk;
          }
        }
        if (k == b) {
          atomicAdd(out ? &to[h] : &ti[h], wt);
          done = true;
        }
        h = (h + 1) & mask;
      }
      full |= !done;
    }
  }
  wave_fence();
  for (int32_t j = lane; j < c; j += kWave) M.kv[o + j] = Slot{tk[j], 0, to[j], ti[j]};
  nk = wave_sum(nk);
  if (lane == 0) M.nkeys[a] = (int32_t)nk;
  if (__ballot(full)) {
    if (lane == 0) atomicOr(flags, kFlagFull);
  }
}

// M build for B <= kSmallB (every row direct): per-workgroup B*B counts in
// LDS, flushed once, so the few hot cells are not hammered by atomics.
__global__ __launch_bounds__(256) void k_insert_small(int64_t E, const int32_t* __restrict__ src,
                                                      const int32_t* __restrict__ dst,
                                                      const int32_t* __restrict__ w,
                                                      const int32_t* __restrict__ labels,
                                                      MView M) {
  __shared__ int32_t s_h[kSmallB * kSmallB];
  const int32_t B = M.B;
  for (int i = threadIdx.x; i < B * B; i += blockDim.x) s_h[i] = 0;
  __syncthreads();
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&s_h[labels[src[e]] * B + labels[dst[e]]], w ? w[e] : 1);
  __syncthreads();
  for (int i = threadIdx.x; i < B * B; i += blockDim.x) {
    const int32_t v = s_h[i];
    if (!v) continue;
    const int32_t a = i / B, b = i - a * B;
    atomicAdd(&M.kv[M.off[a] + b].o, v);
    atomicAdd(&M.kv[M.off[b] + a].i, v);
  }
}

// M build for kSmallB < B <= kPairSortB (rows mostly direct, few distinct
// block pairs, so per-edge atomics would pile onto the same counts): the
// edges' block pairs a*B + b are sorted and summed per pair, then each pair
// is added once.
constexpr int32_t kPairSortB = 4096;
constexpr int32_t kRowBuildB = 16384;  // from here up the build goes by rows (k_build_rows)

__global__ void k_pair_keys(int64_t E, const int32_t* __restrict__ src,
                            const int32_t* __restrict__ dst, const int32_t* __restrict__ w,
                            const int32_t* __restrict__ labels, int32_t B, uint32_t* key,
                            int32_t* val) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    key[e] = (uint32_t)labels[src[e]] * (uint32_t)B + (uint32_t)labels[dst[e]];
    val[e] = w ? w[e] : 1;
  }
}

__global__ void k_insert_pairs(int64_t E, const uint32_t* __restrict__ key,
                               const int32_t* __restrict__ val, const int32_t* count, MView M,
                               int32_t* flags) {
  const int64_t n = min<int64_t>(*count, E);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t a = (int32_t)(key[e] / (uint32_t)M.B), b = (int32_t)(key[e] % (uint32_t)M.B);
    const int32_t w = val[e];
    if (a == b) {
      m_add<false>(M, a, a, w, w, flags);
    } else {
      m_add<false>(M, a, b, w, 0, flags);
      m_add<false>(M, b, a, 0, w, flags);
    }
  }
}

// Blocks holding other than one node (0 means every block is a single node).
__global__ void k_count_nonsingle(int32_t B, const int32_t* __restrict__ bsize, int32_t* out) {
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < B; i += gridDim.x * blockDim.x)
    if (bsize[i] != 1) atomicAdd(out, 1);
}

// Bucket sums (o + i) of every block row (one wavefront per row).
__global__ __launch_bounds__(256) void k_bucket_sums(MView m) {
  __shared__ int32_t s_b[4][kNBK];
  const int w = threadIdx.x / kWave, lane = lane_id();
  const int64_t a = (int64_t)blockIdx.x * 4 + w;
  if (a >= m.B) return;
  int32_t* bk = s_b[w];
  for (int i = lane; i < kNBK; i += kWave) bk[i] = 0;
  wave_fence();
  const int64_t o = m.off[a];
  const int32_t c = m.cap[a];
  const bool direct = c == m.B;
  for (int32_t j = lane; j < c; j += kWave) {
    const Slot q = m.kv[o + j];
    const int32_t k = direct ? j : q.k;
    const int32_t v = q.o + q.i;
    if (k >= 0 && v != 0) atomicAdd(&bk[k >> m.bshift], v);
  }
  wave_fence();
  for (int i = lane; i < kNBK; i += kWave) m.bs[a * kNBK + i] = bk[i];
}

// Workgroup stamps of profiled batch launches (100 MHz wall clock; 0 = the
// workgroup did not run).  Each workgroup of a profiled k_nodal, k_apply or
// k_grow writes its start at t[b] and its end at t[stride + b] with plain
// stores: one atomic per workgroup on a shared word costs more than the
// kernel's own tail.  Per batch slot: k_nodal's starts and ends (G each),
// k_apply's (G + 1 each), k_grow's (kGrowBlocks each).
// The per-block counters and flags a rebuild starts from, zeroed in one
// launch (degrees, edge counts, sizes, key counts, growth marks, parked
// counts: B entries each; build_flags[0..2]).
constexpr int kClearArrays = 8;
struct ClearArgs {
  int32_t* p[kClearArrays];
  int32_t n;
  int32_t* flags;
};

__global__ void k_clear_build(ClearArgs c) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int a = 0; a < kClearArrays; ++a)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < c.n; i += stride)
      c.p[a][i] = 0;
  if (blockIdx.x == 0 && threadIdx.x < 3) c.flags[threadIdx.x] = 0;
}

__global__ void k_stamp_init(unsigned long long* t, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    t[i] = 0ull;
}

// Per batch slot and kernel (block slot*3 + kernel): the first workgroup start
// and the last workgroup end, to out[6*slot + 2*kernel + {0, 1}].
__global__ __launch_bounds__(256) void k_stamp_reduce(const unsigned long long* t, int32_t G,
                                                     int32_t grow_blocks, unsigned long long* out) {
  __shared__ unsigned long long s_lo[256], s_hi[256];
  const int32_t slot = blockIdx.x / 3, k = blockIdx.x % 3;
  const int64_t base = (int64_t)slot * (4 * (int64_t)G + 2 + 2 * grow_blocks) +
                       (k == 0 ? 0 : k == 1 ? 2 * (int64_t)G : 4 * (int64_t)G + 2);
  const int32_t n = k == 0 ? G : k == 1 ? G + 1 : grow_blocks;
  unsigned long long lo = ~0ull, hi = 0ull;
  for (int32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const unsigned long long a = t[base + i], b = t[base + n + i];
    if (a != 0ull) lo = min(lo, a);
    hi = max(hi, b);
  }
  s_lo[threadIdx.x] = lo;
  s_hi[threadIdx.x] = hi;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s_lo[threadIdx.x] = min(s_lo[threadIdx.x], s_lo[threadIdx.x + h]);
      s_hi[threadIdx.x] = max(s_hi[threadIdx.x], s_hi[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[6 * (int64_t)slot + 2 * k] = s_lo[0];
    out[6 * (int64_t)slot + 2 * k + 1] = s_hi[0];
  }
}

// Growth state after a rebuild: nothing parked, the pool used up to
// the new slabs' end (the event counters are kept).
__global__ void k_reset_grow(GrowState* g, unsigned long long used) {
  g->n = 0;
  g->nospace = 0;
  g->ovf_n = 0;
  g->used = used;
}

__global__ void k_relabel(int32_t N, int32_t* labels, const int32_t* __restrict__ map) {
  const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x < N) labels[x] = map[labels[x]];
}

// ---------------------------------------------------------------------------
// Graph ingest on the device: CSR (out-lists) and CSC (in-lists) extended by
// an increment of edges.  Lists keep the reference's neighbour order -- a
// node's neighbours in the order their edges were delivered
// (graph_parallel.hpp:290-294) -- so node x's new list is its old list, then
// its new edges in delivery order.  With cnt[x] new edges at x, the new
// offsets are new_ptr[x] = old_ptr[x] + sum_{y<x} cnt[y]; old edge j of x
// moves by that shift, and the j-th edge of the increment stably sorted by
// x lands at old_ptr[x+1] + j.
// ---------------------------------------------------------------------------

__global__ void k_count_keys(int64_t n, const int32_t* __restrict__ key, int32_t* cnt) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[key[i]], 1);
}

// deg[x] = old list length + new edges of x, deg[N] = 0 (exclusive-scan input).
__global__ void k_new_deg(int32_t N, const int64_t* __restrict__ old_ptr,
                          const int32_t* __restrict__ cnt, int64_t* deg) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < N) deg[x] = old_ptr[x + 1] - old_ptr[x] + cnt[x];
  if (x == N) deg[x] = 0;
}

// Old lists to their new places, one wavefront per node; owner (nullable)
// records the node of each new slot (the CSR's edge sources).
__global__ __launch_bounds__(256) void k_move_lists(int32_t N, const int64_t* __restrict__ old_ptr,
                                                    const int64_t* __restrict__ new_ptr,
                                                    const int32_t* __restrict__ old_idx,
                                                    const int32_t* __restrict__ old_w,
                                                    int32_t* new_idx, int32_t* new_w, int32_t* owner) {
  const int64_t x = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (x >= N) return;
  const int64_t o0 = old_ptr[x], o1 = old_ptr[x + 1], shift = new_ptr[x] - o0;
  for (int64_t j = o0 + lane_id(); j < o1; j += kWave) {
    new_idx[j + shift] = old_idx[j];
    if (new_w) new_w[j + shift] = old_w[j];
    if (owner) owner[j + shift] = (int32_t)x;
  }
}

// The increment's edges (sorted stably by their node key) to their places:
// sorted edge j of node x = key[j] goes to old_ptr[x + 1] + j.  other[] is
// the edge's other end (the target for CSR, the source for CSC).
__global__ void k_place_new(int64_t n, const int32_t* __restrict__ skey,
                            const int32_t* __restrict__ sidx, const int64_t* __restrict__ old_ptr,
                            const int32_t* __restrict__ other, const int32_t* __restrict__ w,
                            int32_t* new_idx, int32_t* new_w, int32_t* owner) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x) {
    const int32_t x = skey[j], i = sidx[j];
    const int64_t pos = old_ptr[x + 1] + j;
    new_idx[pos] = other[i];
    if (new_w) new_w[pos] = w ? w[i] : 1;
    if (owner) owner[pos] = x;
  }
}

__global__ void k_iota(int64_t n, int32_t* v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v[i] = (int32_t)i;
}

// ---------------------------------------------------------------------------
// Entropy, graph_parallel.hpp:1281-1319:
// data_S = -(sum_ab L(M_ab) - sum_a L(d_out[a]) - sum_b L(d_in[b]))
// ---------------------------------------------------------------------------

__device__ void block_store128(i128 v, long* out) {
  __shared__ long s_hi[16];
  __shared__ unsigned long s_lo[16];
  v = wave_sum128(v);
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) {
    s_hi[w] = (long)(v >> 64);
    s_lo[w] = (unsigned long)(u128)v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    i128 t = 0;
    for (int i = 0; i < (int)(blockDim.x / kWave); ++i)
      t += (i128)(((u128)(unsigned long)s_hi[i] << 64) | (u128)s_lo[i]);
    out[0] = (long)(t >> 64);
    out[1] = (long)(unsigned long)(u128)t;
  }
}

__global__ __launch_bounds__(256) void k_entropy(const Slot* __restrict__ kv,
                                                 const GrowState* gst,
                                                 const int32_t* __restrict__ d_out,
                                                 const int32_t* __restrict__ d_in, int32_t B,
                                                 const int64_t* __restrict__ lam, int64_t P,
                                                 long* part) {
  const int64_t used = (int64_t)gst->used;  // every live slab lies below
  const int64_t nslots = used < P ? used : P;
  i128 acc = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += stride) {
    const int32_t v = kv[i].o;  // each M[a][b] once, as o of row a
    if (v > 0) acc += (i128)v * (i128)lam[v];
  }
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < B; b += stride)
    acc -= Lterm(lam, d_out[b]) + Lterm(lam, d_in[b]);
  block_store128(acc, part + 2 * blockIdx.x);
}

// Carry-out order (graph_parallel.hpp:1230) on the device: blocks sorted by
// (dS, block id), dS an int128 ordered as (signed hi, unsigned lo), by two
// stable radix passes -- lo first, then hi with its sign bit flipped.
__global__ void k_sort_init(int32_t B, const unsigned long* __restrict__ lo,
                            unsigned long long* key, int32_t* idx) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < B) {
    key[i] = lo[i];
    idx[i] = i;
  }
}

__global__ void k_sort_hi(int32_t B, const long* __restrict__ hi, const int32_t* __restrict__ perm,
                          unsigned long long* key) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < B) key[i] = (unsigned long long)hi[perm[i]] ^ (1ull << 63);
}

// ---------------------------------------------------------------------------
// Block merge, graph_parallel.hpp:761-988 (+ :349-368)
// ---------------------------------------------------------------------------

struct MergeArgs {
  MView M;
  const int32_t* d_out;
  const int32_t* d_in;
  const int64_t* lam;
  int32_t B;
  int32_t K;
  uint64_t seed, stream;  // pre-drawn mode: stream = the round's first draw
  const uint64_t* uni;    // pre-drawn stream (4 words per draw), else null
  int32_t* best_s;
  long* best_hi;
  unsigned long* best_lo;
};

// L(m1) + L(m2) - L(m1+m2) when both counts are non-zero, else 0.
__device__ __forceinline__ i128 overlap_pair(const int64_t* lam, int32_t m1, int32_t m2) {
  return (m1 > 0 && m2 > 0) ? Lterm(lam, m1) + Lterm(lam, m2) - Lterm(lam, (int64_t)m1 + m2)
                            : (i128)0;
}

// per-lane part of sum_{i not in {r,s}} [L(m1) + L(m2) - L(m1+m2)] over the
// keys where rows r and s of M are both non-zero, and the same over the
// columns (the only terms that survive the factorised merge dS): one pass
// over the shorter block row, probing the other for each of its keys.
__device__ i128 merge_overlap_lane(const MView& m, const int64_t* lam, int32_t r, int32_t s) {
  const int lane = lane_id();
  const bool r_short = m.cap[r] <= m.cap[s];
  const int32_t a1 = r_short ? r : s, a2 = r_short ? s : r;
  const Row R1 = m_row(m, a1), R2 = m_row(m, a2);
  const bool direct = R1.c == m.B;
  i128 acc = 0;
  if (direct) {  // then so is the longer row: key j sits at slot j of both
    for (int32_t j = lane; j < R1.c; j += kWave) {
      const int2 c1 = *reinterpret_cast<const int2*>(&m.kv[R1.o + j].o);
      const int2 c2 = *reinterpret_cast<const int2*>(&m.kv[R2.o + j].o);
      if (j == r || j == s) continue;
      acc += overlap_pair(lam, c1.x, c2.x) + overlap_pair(lam, c1.y, c2.y);
    }
    return acc;
  }
  // hashed shorter row: kOvl of its slots per lane, all their probes of
  // the other row issued before any is resolved
  constexpr int kOvl = 4;
  for (int32_t base = 0; base < R1.c; base += kOvl * kWave) {
    Slot q[kOvl];
#pragma unroll
    for (int u = 0; u < kOvl; ++u) {
      const int32_t j = base + u * kWave + lane;
      q[u] = j < R1.c ? m.kv[R1.o + j] : Slot{-1, 0, 0, 0};
      if (q[u].k < 0 || (q[u].o <= 0 && q[u].i <= 0) || q[u].k == r || q[u].k == s) q[u].k = -1;
    }
    Probe p[kOvl];
#pragma unroll
    for (int u = 0; u < kOvl; ++u)
      if (q[u].k >= 0) p[u] = m_probe(m, R2, q[u].k);
#pragma unroll
    for (int u = 0; u < kOvl; ++u) {
      if (q[u].k < 0) continue;
      const Cnt c2 = m_value(m, R2, q[u].k, p[u]);
      acc += overlap_pair(lam, q[u].o, c2.o) + overlap_pair(lam, q[u].i, c2.i);
    }
  }
  return acc;
}

// Row r of a k_merge workgroup in LDS, shared by its K proposals: its
// non-zero keys in ascending order with both counts and the running sum of
// o + i (n = -1: the row has more than kRowLds non-zero keys and the
// proposals read it from M).
constexpr int kRowLds = 256;
struct RowLds {
  const int32_t* k;
  const int32_t* o;
  const int32_t* i;
  const long* cum;
  int32_t n;
};

// Index of the first entry whose running sum exceeds t (t < the row's total).
__device__ __forceinline__ int32_t rowlds_pick(const RowLds& L, int64_t t) {
  int32_t lo = 0, hi = L.n - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (L.cum[mid] > t)
      hi = mid;
    else
      lo = mid + 1;
  }
  return L.k[lo];
}

// Position of key b among the row's keys, or -1.
__device__ __forceinline__ int32_t rowlds_find(const RowLds& L, int32_t b) {
  int32_t lo = 0, hi = L.n;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (L.k[mid] < b)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < L.n && L.k[lo] == b ? lo : -1;
}

// merge_overlap_lane with row r from LDS: r's entries probe row s, or, when
// s's slab is at most four times r's key count, s's slab is scanned and its
// keys are looked up in r.
__device__ i128 merge_overlap_lds(const MView& m, const int64_t* lam, int32_t r, int32_t s,
                                  const RowLds& L) {
  const int lane = lane_id();
  const Row RS = m_row(m, s);
  i128 acc = 0;
  if (RS.c <= 4 * L.n) {
    const bool direct = RS.c == m.B;
    for (int32_t j = lane; j < RS.c; j += kWave) {
      const Slot q = m.kv[RS.o + j];
      const int32_t i = direct ? j : q.k;
      if (i < 0 || (q.o <= 0 && q.i <= 0) || i == r || i == s) continue;
      const int32_t p = rowlds_find(L, i);
      if (p < 0) continue;
      const int32_t o1 = L.o[p], i1 = L.i[p];
      if (o1 > 0 && q.o > 0) acc += Lterm(lam, o1) + Lterm(lam, q.o) - Lterm(lam, (int64_t)o1 + q.o);
      if (i1 > 0 && q.i > 0) acc += Lterm(lam, i1) + Lterm(lam, q.i) - Lterm(lam, (int64_t)i1 + q.i);
    }
  } else {
    for (int32_t e = lane; e < L.n; e += kWave) {
      const int32_t i = L.k[e];
      if (i == r || i == s) continue;
      const Cnt c2 = m_value(m, RS, i, m_probe(m, RS, i));
      const int32_t o1 = L.o[e], i1 = L.i[e];
      if (o1 > 0 && c2.o > 0) acc += Lterm(lam, o1) + Lterm(lam, c2.o) - Lterm(lam, (int64_t)o1 + c2.o);
      if (i1 > 0 && c2.i > 0) acc += Lterm(lam, i1) + Lterm(lam, c2.i) - Lterm(lam, (int64_t)i1 + c2.i);
    }
  }
  return acc;
}

__device__ void merge_one(const MergeArgs& a, int32_t r, int32_t idx, int32_t* scr,
                          const RowLds& L, int32_t& s_out, i128& dS_out) {
  const int32_t B = a.B;
  const int64_t dor = a.d_out[r], dir = a.d_in[r];
  const int64_t k = dor + dir;
  uint64_t x[4];
  if (a.uni)
    predrawn(a.uni, a.stream + (uint64_t)r * (uint64_t)a.K + (uint64_t)idx, x);
  else
    usap::philox4x64((uint64_t)r, (uint64_t)idx, a.stream, usap::kTagMerge, a.seed, 0, x);
  int32_t s;
  if (k == 0) {
    s = (int32_t)usap::idx_of(x[2], B);
  } else {
    long bytes = 0;
    const int64_t t0 = (int64_t)usap::idx_of(x[0], k);
    const int32_t u = L.n >= 0 ? rowlds_pick(L, t0) : select_combined(a.M, r, t0, -1, scr, kScr, bytes);
    const int64_t da = (int64_t)a.d_out[u] + a.d_in[u];
    if (usap::uniform_branch(x[1], da, B)) {
      s = usap::uniform_not_r(x[2], r, B);
    } else {
      const Cnt ur = m_get(a.M, u, r);  // M[u][r], M[r][u]
      const int64_t W2 = da - ur.o - ur.i;
      s = W2 == 0 ? usap::uniform_not_r(x[2], r, B)
                  : select_combined(a.M, u, usap::idx_of(x[3], W2), r, scr, kScr, bytes);
    }
  }
  s_out = s;
  if (s == r) {
    dS_out = 0;
    return;
  }
  const int64_t* lam = a.lam;
  const Cnt crs = m_get(a.M, r, s);  // M[r][s], M[s][r]
  const int64_t Mrr = m_get(a.M, r, r).o, Mrs = crs.o;
  const int64_t Msr = crs.i, Mss = m_get(a.M, s, s).o;
  const int64_t dos = a.d_out[s], dis = a.d_in[s];
  const int64_t dos2 = dos + dor, dis2 = dis + dir;
  i128 acc = L.n >= 0 ? merge_overlap_lds(a.M, lam, r, s, L) : merge_overlap_lane(a.M, lam, r, s);
  acc = wave_sum128(acc);
  bool inf = false;
  const i128 old_c = Tterm(lam, Mrr, dir, dor, inf) + Tterm(lam, Mrs, dis, dor, inf) +
                     Tterm(lam, Msr, dir, dos, inf) + Tterm(lam, Mss, dis, dos, inf);
  const i128 new_c = Tterm(lam, Mrr + Mrs + Msr + Mss, dis2, dos2, inf);
  const int64_t Xr = dor - Mrr - Mrs, Xs = dos - Msr - Mss;
  const int64_t Yr = dir - Mrr - Msr, Ys = dis - Mrs - Mss;
  const i128 deg = -Pterm(lam, Xr, dor, inf) - Pterm(lam, Xs, dos, inf) +
                   Pterm(lam, Xr + Xs, dos2, inf) - Pterm(lam, Yr, dir, inf) -
                   Pterm(lam, Ys, dis, inf) + Pterm(lam, Yr + Ys, dis2, inf);
  dS_out = old_c - new_c + acc + deg;
}

__global__ __launch_bounds__(kWave* kMergeWaves, 5) void k_merge(MergeArgs a) {
  __shared__ long s_hi[kMergeWaves];
  __shared__ unsigned long s_lo[kMergeWaves];
  __shared__ int32_t s_idx[kMergeWaves];
  __shared__ int32_t s_s[kMergeWaves];
  __shared__ int32_t s_scr[kMergeWaves][kScr];
  __shared__ int32_t s_rk[kRowLds], s_ro[kRowLds], s_ri[kRowLds];
  __shared__ long s_cum[kRowLds];
  __shared__ int32_t s_rn;
  const int32_t r = blockIdx.x;
  const int w = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  // row r's non-zero keys into LDS once for the K proposals (unordered), then
  // ordered by key (ranks: keys are distinct), then the running sums
  const Row R = m_row(a.M, r);
  const bool fits = R.c <= 8 * kRowLds;  // the same for every thread
  if (threadIdx.x == 0) s_rn = 0;
  __syncthreads();
  if (fits) {
    const bool direct = R.c == a.B;
    for (int32_t j = threadIdx.x; j < R.c; j += blockDim.x) {
      const Slot q = a.M.kv[R.o + j];
      const int32_t key = direct ? j : q.k;
      if (key >= 0 && (q.o > 0 || q.i > 0)) {
        const int32_t p = atomicAdd(&s_rn, 1);
        if (p < kRowLds) {
          s_rk[p] = key;
          s_ro[p] = q.o;
          s_ri[p] = q.i;
        }
      }
    }
  }
  __syncthreads();
  const int32_t rn = fits ? s_rn : kRowLds + 1;
  RowLds L{s_rk, s_ro, s_ri, s_cum, rn <= kRowLds ? rn : -1};
  if (L.n > 0) {
    constexpr int kPer = kRowLds / kWave;  // entries per thread (blockDim >= one wavefront)
    int32_t mk[kPer], mo[kPer], mi[kPer], rank[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int32_t e = threadIdx.x + q * blockDim.x;
      rank[q] = -1;
      if (e < L.n) {
        mk[q] = s_rk[e];
        mo[q] = s_ro[e];
        mi[q] = s_ri[e];
        rank[q] = 0;
        for (int32_t j = 0; j < L.n; ++j) rank[q] += s_rk[j] < mk[q];
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      if (rank[q] < 0) continue;
      s_rk[rank[q]] = mk[q];
      s_ro[rank[q]] = mo[q];
      s_ri[rank[q]] = mi[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      long c = 0;
      for (int32_t j = 0; j < L.n; ++j) {
        c += (long)s_ro[j] + s_ri[j];
        s_cum[j] = c;
      }
    }
    __syncthreads();
  }
  int32_t bidx = -1, bs = -1;
  i128 bd = 0;
  for (int32_t idx = w; idx < a.K; idx += nw) {
    int32_t s;
    i128 d;
    merge_one(a, r, idx, s_scr[w], L, s, d);
    if (bidx < 0 || d < bd) {  // strict less: the earliest proposal wins ties
      bidx = idx;
      bs = s;
      bd = d;
    }
  }
  if (lane_id() == 0) {
    s_hi[w] = (long)(bd >> 64);
    s_lo[w] = (unsigned long)(u128)bd;
    s_idx[w] = bidx;
    s_s[w] = bs;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int best = -1;
    i128 bv = 0;
    for (int i = 0; i < nw; ++i) {
      if (s_idx[i] < 0) continue;
      const i128 v = (i128)(((u128)(unsigned long)s_hi[i] << 64) | (u128)s_lo[i]);
      if (best < 0 || v < bv || (v == bv && s_idx[i] < s_idx[best])) {
        best = i;
        bv = v;
      }
    }
    a.best_s[r] = s_s[best];
    a.best_hi[r] = s_hi[best];
    a.best_lo[r] = s_lo[best];
  }
}

// ---------------------------------------------------------------------------
// Nodal proposal, graph_parallel.hpp:990-1222
// ---------------------------------------------------------------------------

constexpr int kEdgeRegs = 4;     // list entries per lane per pass (256 per wavefront)
constexpr int kScrNodal = 1024;  // k_nodal's LDS scratch per wavefront (one key bucket)

// Workgroups are dispatched to the GPU's 8 XCDs (each with its own L2) in
// turn: workgroup b runs on XCD b % 8.  Block b of a batch kernel takes item
// block xcd_block(b, G) instead of b, so the workgroups of one XCD cover
// consecutive nodes, and the cache lines holding those nodes' CSR/CSC offsets,
// lists and proposal outputs are fetched into one L2 instead of up to eight.
// A bijection of [0, G).
constexpr int kXcds = 8;
__device__ __forceinline__ int32_t xcd_block(int32_t b, int32_t G) {
  const int32_t per = G / kXcds, extra = G % kXcds;
  const int32_t x = b % kXcds, j = b / kXcds;
  return (x < extra ? x * (per + 1) : extra * (per + 1) + (x - extra) * per) + j;
}

// One workgroup's share of a batch, written by k_nodal and summed by block 0
// of k_apply (no same-address atomics per node).
struct BlockPart {
  long sum_hi;  // sum of the accepted finite dS of the block's nodes (int128)
  unsigned long sum_lo;
  unsigned long long nodal_bytes;  // algorithmic bytes of k_nodal
  unsigned long long apply_bytes;  // algorithmic bytes of k_apply
  int32_t moves;
  int32_t ninf;
};

// Algorithmic bytes of one accepted move of a node with kc list entries: the
// list (id, label, weight), per entry up to 4 slots of 12 bytes, the rows
// and degrees of r and s, the label.
__device__ __forceinline__ long apply_tally(int64_t kc) { return 52 + 60 * kc + 32 + 32 + 4; }

struct NodalArgs {
  GraphView g;
  MView M;
  const int32_t* labels;
  const int32_t* d_out;
  const int32_t* d_in;
  const int64_t* lam;
  int32_t B;
  int32_t beg, end;
  int32_t exact;
  int32_t bidx;  // batch index in the sweep: random stream sweep->stream0 + bidx
  uint64_t seed;
  const uint64_t* uni;  // pre-drawn stream: node n takes draw sweep->stream0 + n; else null
  const int64_t* big_off;
  int32_t* big_tab;
  int32_t* prop_s;
  long* prop_hi;
  unsigned long* prop_lo;
  int32_t* prop_inf;
  int32_t* prop_acc;
  int32_t* prop_r;
  int32_t* prop_u;  // profiling: block of the drawn neighbour (-1: none), else null
  BlockPart* part;
  Status* status;
  Sweep* sweep;  // read, and stop_n written by block 0
  // kernel timers on (profiling): [0] earliest workgroup start, [1] latest
  // workgroup end of this launch, 100 MHz clock; else null
  unsigned long long* tstamp;  // workgroup stamps (start [b], end [tstride + b]) or null
  int32_t tstride;
};

__device__ __forceinline__ void tab_add(int32_t* tk, int32_t* tv, uint32_t mask, int32_t b,
                                        int32_t w) {
  uint32_t h = hash32(b) & mask;
  for (;;) {
    int32_t k = __hip_atomic_load(&tk[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    if (k < 0) {
      k = atomicCAS(&tk[h], -1, b);
      if (k == -1) k = b;
    }
    if (k == b) {
      atomicAdd(&tv[h], w);
      return;
    }
    h = (h + 1) & mask;
  }
}

// The neighbour-block table of node n: slots [0, cap) of tk (block), two
// (weight of n's out-edges into it) and twi (in-edges), in LDS for nodes with
// at most kTabMaxDeg list entries, else in the node's global-memory table.
struct Table {
  int32_t *tk, *two, *twi;
  int32_t cap;
};

// kBig selects the node's global table (more than kTabMaxDeg list entries)
// or the wavefront's LDS table at compile time, so the LDS form compiles to
// LDS instructions (a pointer that may be either would make every table
// access a flat access that also waits for the wavefront's global loads).
template <bool kBig>
__device__ __forceinline__ Table tab_open(int32_t* lds, const int64_t* big_off, int32_t* big_tab,
                                          int32_t n, int64_t kc) {
  const int lane = lane_id();
  Table T;
  if (kBig) {
    int32_t cap = 1;
    while (cap < 2 * kc) cap <<= 1;
    T.cap = cap;
    T.tk = big_tab + big_off[n];
    T.two = T.tk + cap;
    T.twi = T.tk + 2 * cap;
  } else {
    int32_t cap = kWave;  // >= 4/3 of the distinct neighbour blocks, <= kTab
    while (cap < 2 * kc && cap < kTab) cap <<= 1;
    T.cap = cap;
    T.tk = lds;
    T.two = lds + kTab;
    T.twi = lds + 2 * kTab;
  }
  for (int32_t i = lane; i < T.cap; i += kWave) {
    T.tk[i] = -1;
    T.two[i] = 0;
    T.twi[i] = 0;
  }
  wave_fence();
  return T;
}

// Moves the table's entries, except those of blocks r and s (whose weights
// are returned in wor/wir/wos/wis), to its front in place; returns their
// number.  Chunk by chunk: every lane reads its slot before any write, and a
// write never lands after the chunk being read.
__device__ __forceinline__ int32_t tab_compact(const Table& T, int32_t r, int32_t s, long& wor,
                                               long& wir, long& wos, long& wis) {
  const int lane = lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  int32_t nent = 0;
  for (int32_t base = 0; base < T.cap; base += kWave) {  // cap is a multiple of kWave
    const int32_t i = base + lane;
    int32_t b = T.tk[i];
    const int32_t wo = T.two[i], wi = T.twi[i];
    const bool isr = b >= 0 && b == r, iss = b >= 0 && b == s;
    wor += isr ? wo : 0;
    wir += isr ? wi : 0;
    wos += iss ? wo : 0;
    wis += iss ? wi : 0;
    if (isr || iss) b = -1;
    const unsigned long long live = __ballot(b >= 0);
    const int32_t pos = nent + __popcll(live & below);
    wave_fence();
    if (b >= 0) {
      T.tk[pos] = b;
      T.two[pos] = wo;
      T.twi[pos] = wi;
    }
    nent += __popcll(live);
    wave_fence();
  }
  return nent;
}

// List entry j of node n (out-edges then in-edges): neighbour id (-1 past the
// end) and weight.
__device__ __forceinline__ void list_entry(const GraphView& g, int64_t o0, int64_t kco,
                                           int64_t i0, int64_t kc, int64_t j, int32_t& v,
                                           int32_t& w) {
  v = -1;
  w = 1;
  if (j < kco) {
    v = g.out_idx[o0 + j];
    if (g.out_w) w = g.out_w[o0 + j];
  } else if (j < kc) {
    v = g.in_idx[i0 + (j - kco)];
    if (g.in_w) w = g.in_w[i0 + (j - kco)];
  }
}

// dS of moving node n from r to s (r != s), factorised, from its complete
// neighbour-block table.  All lookups of a pass are issued before any is
// resolved.
__device__ __forceinline__ i128 nodal_delta(const NodalArgs& a, const Table& T, int32_t r,
                                            int32_t s, Row rM, int64_t dor, int64_t dir,
                                            int64_t kc, int64_t k_out, int64_t k_in, long wself,
                                            bool& inf, long& bytes, int32_t s_pre, Row sM_pre,
                                            int64_t dos_pre, int64_t dis_pre) {
  const int lane = lane_id();
  const int64_t* lam = a.lam;
  // s's row and degrees (loaded ahead when s is the uniform draw); the table
  // is compacted while they load
  const bool pre = s == s_pre;
  const Row sM = pre ? sM_pre : m_row(a.M, s);
  const int64_t dos = pre ? dos_pre : a.d_out[s], dis = pre ? dis_pre : a.d_in[s];
  long wor = 0, wir = 0, wos = 0, wis = 0;
  const int32_t nent = tab_compact(T, r, s, wor, wir, wos, wis);
  // corner: (r, r), (r, s) with (s, r) beside it, (s, s)
  const Probe pRR = m_probe(a.M, rM, r), pRS = m_probe(a.M, rM, s);
  const Probe pSS = m_probe(a.M, sM, s);
  // changed entries of rows r, s (out-edges, o) and columns r, s (in-edges,
  // i): one probe of block row r and one of s per neighbour block
  i128 acc = 0;
  long lookups = 0;
  bool neg = false;
  for (int32_t j = lane; j < nent; j += kWave) {
    const int32_t b = T.tk[j], wo = T.two[j], wi = T.twi[j];
    const Probe p0 = m_probe(a.M, rM, b), p1 = m_probe(a.M, sM, b);
    const Cnt cr = m_value(a.M, rM, b, p0), cs = m_value(a.M, sM, b, p1);
    lookups += 2;
    if (wo > 0) {
      const int64_t Rr = cr.o, Rs = cs.o;
      neg |= Rr - wo < 0;
      acc += Lterm(lam, Rr) + Lterm(lam, Rs) - Lterm(lam, Rr - wo) - Lterm(lam, Rs + wo);
    }
    if (wi > 0) {
      const int64_t Cr = cr.i, Cs = cs.i;
      neg |= Cr - wi < 0;
      acc += Lterm(lam, Cr) + Lterm(lam, Cs) - Lterm(lam, Cr - wi) - Lterm(lam, Cs + wi);
    }
  }
  if (neg) atomicOr(&a.status->flags, kFlagNegative);
  acc = wave_sum128(acc);
  // blocks r and s own one table slot each, so one lane at most holds their
  // weights; the other entries' weights are the rest of k_out / k_in
  const unsigned long long hr = __ballot(wor != 0 || wir != 0);
  const unsigned long long hs = __ballot(wos != 0 || wis != 0);
  wor = hr ? read_lane(wor, __ffsll(hr) - 1) : 0;
  wir = hr ? read_lane(wir, __ffsll(hr) - 1) : 0;
  wos = hs ? read_lane(wos, __ffsll(hs) - 1) : 0;
  wis = hs ? read_lane(wis, __ffsll(hs) - 1) : 0;
  const long Wo = k_out - wor - wos, Wi = k_in - wir - wis;
  if (__ballot(wself != 0)) wself = wave_sum(wself);  // self-loops are rare
  // neighbour ids + labels (+ weights); rows (offset + capacity) and degrees
  // of r and s; 2 block-row lookups (key + two counts) per changed entry and
  // 3 corner lookups
  bytes += (a.g.out_w ? 12 : 8) * kc + 24 + 16 + 12 * (wave_sum(lookups) + 3);
  const Cnt crs = m_value(a.M, rM, s, pRS);
  const int64_t Mrr = m_value(a.M, rM, r, pRR).o, Mrs = crs.o;
  const int64_t Msr = crs.i, Mss = m_value(a.M, sM, s, pSS).o;
  const int64_t dor2 = dor - k_out, dos2 = dos + k_out, dir2 = dir - k_in, dis2 = dis + k_in;
  int64_t rr2, rs2, sr2, ss2;
  if (!a.exact) {  // graph_parallel.hpp:1117-1145: only out-edges update the rows
    rr2 = Mrr - wor;
    rs2 = Mrs - wos;
    sr2 = Msr + wor;
    ss2 = Mss + wos;
  } else {  // graph.cu:1414-1426, with a self-loop moving (r,r) -> (s,s)
    const int64_t wor_ = wor - wself, wir_ = wir - wself;
    rr2 = Mrr - wor_ - wself - wir_;
    rs2 = Mrs - wos + wir_;
    sr2 = Msr + wor_ - wis;
    ss2 = Mss + wos + wis + wself;
  }
  bool old_inf = false;
  const i128 old_c = Tterm(lam, Mrr, dir, dor, old_inf) + Tterm(lam, Mrs, dis, dor, old_inf) +
                     Tterm(lam, Msr, dir, dos, old_inf) + Tterm(lam, Mss, dis, dos, old_inf);
  const i128 new_c = Tterm(lam, rr2, dir2, dor2, inf) + Tterm(lam, rs2, dis2, dor2, inf) +
                     Tterm(lam, sr2, dir2, dos2, inf) + Tterm(lam, ss2, dis2, dos2, inf);
  const int64_t Xr = dor - Mrr - Mrs, Xs = dos - Msr - Mss;
  const int64_t Yr = dir - Mrr - Msr, Ys = dis - Mrs - Mss;
  const i128 deg_old = Pterm(lam, Xr, dor, old_inf) + Pterm(lam, Xs, dos, old_inf) +
                       Pterm(lam, Yr, dir, old_inf) + Pterm(lam, Ys, dis, old_inf);
  const i128 deg_new = Pterm(lam, Xr - Wo, dor2, inf) + Pterm(lam, Xs + Wo, dos2, inf) +
                       Pterm(lam, Yr - Wi, dir2, inf) + Pterm(lam, Ys + Wi, dis2, inf);
  if (rr2 < 0 || rs2 < 0 || sr2 < 0 || ss2 < 0 || Xr - Wo < 0 || Yr - Wi < 0)
    atomicOr(&a.status->flags, kFlagNegative);
  return old_c - new_c + acc - deg_old + deg_new;
}

// One nodal proposal (graph_parallel.hpp:990-1222) for node n, batch item
// `item`; returns its algorithmic bytes.  The loads form a short chain: own
// label and list bounds; the list (ids) with r's rows and degrees; the
// neighbours' labels with the drawn neighbour's; its block's degrees (and, when
// the proposal follows M, its rows' bucket sums and one bucket); s's rows and
// degrees; then every lookup of the dS at once.
template <bool kBig>
__device__ __forceinline__ long nodal_one(const NodalArgs& a, int64_t item, int32_t n, int32_t r,
                                          int64_t o0, int64_t kco, int64_t i0, int64_t kci,
                                          uint64_t stream0, int32_t* tab, int32_t* scr,
                                          int32_t& acc_out, int32_t& inf_out, i128& dS_out) {
  const int lane = lane_id();
  const GraphView& g = a.g;
  const int32_t B = a.B;
  const int64_t kc = kco + kci;
  uint64_t x[4];
  if (a.uni)
    predrawn(a.uni, stream0 + (uint64_t)n, x);
  else
    usap::philox4x64((uint64_t)n, 0, stream0 + (uint64_t)a.bidx, usap::kTagNodal, a.seed, 0, x);
  const Row rM = m_row(a.M, r);
  const int64_t dor = a.d_out[r], dir = a.d_in[r];
  // the uniform draw's block, the proposal in most cases: its row and
  // degrees load with r's
  const int32_t s_pre = (int32_t)usap::idx_of(x[2], B);
  const Row sM_pre = m_row(a.M, s_pre);
  const int64_t dos_pre = a.d_out[s_pre], dis_pre = a.d_in[s_pre];
  int64_t k_out = kco, k_in = kci;
  if (g.out_w) {
    long so = 0, si = 0;
    for (int64_t j = lane; j < kco; j += kWave) so += g.out_w[o0 + j];
    for (int64_t j = lane; j < kci; j += kWave) si += g.in_w[i0 + j];
    k_out = wave_sum(so);
    k_in = wave_sum(si);
  }
  const int64_t k = k_out + k_in;
  // list offsets, own label, outputs (s, dS, inf, accept, r) (+ weights)
  long bytes = 32 + 4 + 32 + (g.out_w ? 4 * kc : 0);
  const Table T = tab_open<kBig>(tab, a.big_off, a.big_tab, n, kc);
  const uint32_t mask = (uint32_t)T.cap - 1;
  // the drawn neighbour (graph_parallel.hpp:1010-1027): unit t of the list
  const int64_t t = k > 0 ? (int64_t)usap::idx_of(x[0], k) : 0;
  int32_t vpick = -1;
  if (k > 0) {
    if (g.out_w)
      vpick = pick_neighbor(g, o0, kco, i0, kci, t);
    else if (t >= kEdgeRegs * kWave)
      vpick = t < kco ? g.out_idx[o0 + t] : g.in_idx[i0 + (t - kco)];
  }
  int32_t u = -1;
  int64_t da = 0;
  long wself = 0;
  for (int64_t base = 0; base < kc; base += kEdgeRegs * kWave) {
    int32_t v[kEdgeRegs], ww[kEdgeRegs], lb[kEdgeRegs];
#pragma unroll
    for (int q = 0; q < kEdgeRegs; ++q)
      list_entry(g, o0, kco, i0, kc, base + q * kWave + lane, v[q], ww[q]);
    if (base == 0 && k > 0 && vpick < 0) {  // unweighted, t < 256: lane t % 64 holds it
      int32_t vv = v[0];
#pragma unroll
      for (int q = 1; q < kEdgeRegs; ++q)
        if ((t >> 6) == q) vv = v[q];
      vpick = __shfl(vv, (int)(t & (kWave - 1)));
    }
#pragma unroll
    for (int q = 0; q < kEdgeRegs; ++q) lb[q] = v[q] >= 0 ? a.labels[v[q]] : -1;
    if (base == 0 && k > 0) {
      u = a.labels[vpick];
      da = (int64_t)a.d_out[u] + a.d_in[u];
    }
#pragma unroll
    for (int q = 0; q < kEdgeRegs; ++q) {
      if (v[q] < 0) continue;
      if (base + q * kWave + lane < kco) {
        if (v[q] == n) wself += ww[q];
        tab_add(T.tk, T.two, mask, lb[q], ww[q]);
      } else {
        tab_add(T.tk, T.twi, mask, lb[q], ww[q]);
      }
    }
  }
  int32_t s;
  if (k == 0) {
    s = (int32_t)usap::idx_of(x[2], B);
  } else {
    bytes += 16;  // drawn neighbour id, its label, d_out[u], d_in[u]
    if (usap::uniform_branch(x[1], da, B))
      s = (int32_t)usap::idx_of(x[2], B);
    else
      s = select_combined(a.M, u, usap::idx_of(x[3], da), -1, scr, kScrNodal, bytes);
  }
  i128 dS = 0;
  bool inf = false;
  int32_t acc = 0;
  if (s != r) {
    wave_fence();  // the table is complete
    dS = nodal_delta(a, T, r, s, rM, dor, dir, kc, k_out, k_in, wself, inf, bytes, s_pre, sM_pre,
                     dos_pre, dis_pre);
    acc = (inf || dS < 0) ? 1 : 0;
  }
  if (lane == 0) {
    a.prop_s[item] = s;
    a.prop_hi[item] = (long)(dS >> 64);
    a.prop_lo[item] = (unsigned long)(u128)dS;
    a.prop_inf[item] = inf ? -1 : 0;
    a.prop_acc[item] = acc;
    a.prop_r[item] = r;
    if (a.prop_u) a.prop_u[item] = u;
  }
  acc_out = acc;
  inf_out = inf ? 1 : 0;
  dS_out = dS;
  return bytes;
}

__global__ __launch_bounds__(kWave* kNodalWaves) void k_nodal(NodalArgs a) {
  __shared__ int32_t s_tab[kNodalWaves][3 * kTab];
  __shared__ __align__(8) int32_t s_scr[kNodalWaves][kScrNodal];
  const int w = threadIdx.x / kWave;
  if (a.tstamp && threadIdx.x == 0) a.tstamp[blockIdx.x] = wall_clock64();
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.M.gst->ovf_n = 0;  // the previous k_grow replayed the parked increments
    // k_apply decides by this copy whether the batch runs: its block 0
    // updates stop meanwhile
    a.sweep->stop_n = a.sweep->stop;
  }
  const int64_t item = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kNodalWaves + w;
  const int64_t n64 = (int64_t)a.beg + item;
  long bytes = 0, abytes = 0;
  int32_t acc = 0, inf = 0;
  i128 dS = 0;
  if (n64 < a.end) {
    // the sweep's stop flag is read together with the node's first loads
    const int32_t n = (int32_t)n64;
    const int32_t stop = a.sweep->stop;
    const uint64_t stream0 = a.sweep->stream0;
    const int32_t r = a.labels[n];
    const int64_t o0 = a.g.out_ptr[n], i0 = a.g.in_ptr[n];
    const int64_t kco = a.g.out_ptr[n + 1] - o0, kci = a.g.in_ptr[n + 1] - i0;
    if (!stop) {
      bytes = kco + kci > kTabMaxDeg
                  ? nodal_one<true>(a, item, n, r, o0, kco, i0, kci, stream0, s_tab[w], s_scr[w],
                                    acc, inf, dS)
                  : nodal_one<false>(a, item, n, r, o0, kco, i0, kci, stream0, s_tab[w],
                                     s_scr[w], acc, inf, dS);
      if (acc) abytes = apply_tally(kco + kci);
    }
  }
  __syncthreads();  // every wavefront is done with its scratch
  // this block's share of the batch sums: accepted moves, the exact sum of
  // their finite dS, the -inf ones, byte tallies
  unsigned long long* s_p = reinterpret_cast<unsigned long long*>(&s_scr[0][0]);
  if (lane_id() == 0) {
    const u128 d = (acc && !inf) ? (u128)dS : (u128)0;
    s_p[6 * w + 0] = (unsigned long long)bytes;
    s_p[6 * w + 1] = (unsigned long long)abytes;
    s_p[6 * w + 2] = (unsigned long long)d;
    s_p[6 * w + 3] = (unsigned long long)(d >> 64);
    s_p[6 * w + 4] = (unsigned long long)acc;
    s_p[6 * w + 5] = (unsigned long long)(acc && inf);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long nb = 0, ab = 0;
    u128 sum = 0;
    int32_t moves = 0, ninf = 0;
    for (int i = 0; i < kNodalWaves; ++i) {
      nb += s_p[6 * i + 0];
      ab += s_p[6 * i + 1];
      sum += ((u128)s_p[6 * i + 3] << 64) | (u128)s_p[6 * i + 2];
      moves += (int32_t)s_p[6 * i + 4];
      ninf += (int32_t)s_p[6 * i + 5];
    }
    BlockPart& P = a.part[blockIdx.x];
    P.sum_lo = (unsigned long)sum;
    P.sum_hi = (long)(sum >> 64);
    P.nodal_bytes = nb;
    P.apply_bytes = ab;
    P.moves = moves;
    P.ninf = ninf;
    if (a.tstamp) a.tstamp[a.tstride + blockIdx.x] = wall_clock64();
  }
}

// Profiling only: the compulsory bytes of the reference's algorithm for one
// nodal batch, from the batch's own proposals.  graph_parallel.hpp:990-1222
// reads, per node n, its list offsets (16), label (4), neighbour ids and
// labels (8 k_n, + 4 k_n weights) and writes s and dS (12); per distinct
// block b among the batch's old blocks r_n, proposed blocks s_n and drawn
// neighbours' blocks u_n, row and column b of M whole (16 + 8 nnzR(b) +
// 8 nnzC(b)) and its degrees (12); per accepted move, the label (4) and two
// 8-byte read-modify-writes per list entry (16 k_n).  A block counts once per
// batch: the first wavefront to raise its mark to the batch's stamp counts
// the non-zeros of its block row (o > 0: row b, i > 0: column b).  One
// wavefront per batch node; each workgroup adds its per-node bytes to
// acc[2 j] and its block bytes to acc[2 j + 1], j = workgroup % kAqSlots
// (spread over slots: one address would serialise the whole grid's adds).
constexpr int kAqSlots = 64;
struct AqArgs {
  GraphView g;
  MView M;
  int32_t beg, end;
  const int32_t* prop_r;
  const int32_t* prop_s;
  const int32_t* prop_u;
  const int32_t* prop_acc;
  int32_t* mark;
  int32_t stamp;
  const Sweep* sweep;
  unsigned long long* acc;
};

__global__ __launch_bounds__(kWave* kNodalWaves) void k_aq(AqArgs a) {
  const int lane = lane_id();
  const int64_t item = (int64_t)blockIdx.x * kNodalWaves + threadIdx.x / kWave;
  const int64_t n64 = (int64_t)a.beg + item;
  const bool live = n64 < a.end && a.sweep->stop_n == 0;  // uniform per wavefront
  const int32_t n = live ? (int32_t)n64 : a.beg;
  const int64_t kc =
      live ? (a.g.out_ptr[n + 1] - a.g.out_ptr[n]) + (a.g.in_ptr[n + 1] - a.g.in_ptr[n]) : 0;
  int32_t b = -1;
  if (live && lane == 0) b = a.prop_r[item];
  if (live && lane == 1) b = a.prop_s[item];
  if (live && lane == 2) b = a.prop_u[item];
  const bool first = b >= 0 && atomicMax(&a.mark[b], a.stamp) < a.stamp;
  unsigned long long rows = __ballot(first);
  unsigned long long blk = 0;
  while (rows) {
    const int l = __ffsll(rows) - 1;
    rows &= rows - 1ull;
    const int32_t bb = __shfl(b, l);
    const int64_t o = a.M.off[bb];
    const int32_t c = a.M.cap[bb];
    long nz = 0;
    for (int32_t j = lane; j < c; j += kWave) {
      const Slot q = a.M.kv[o + j];
      nz += (q.o > 0 ? 1 : 0) + (q.i > 0 ? 1 : 0);
    }
    blk += 16 + 8 * (unsigned long long)wave_sum(nz) + 12;
  }
  __shared__ unsigned long long s_node[kNodalWaves], s_blk[kNodalWaves];
  const int w = threadIdx.x / kWave;
  if (lane == 0) {
    unsigned long long node = 0;
    if (live) {
      node = 16 + 4 + 8 * (unsigned long long)kc + 12;
      if (a.g.out_w) node += 4 * (unsigned long long)kc;
      if (a.prop_acc[item]) node += 4 + 16 * (unsigned long long)kc;
    }
    s_node[w] = node;
    s_blk[w] = blk;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long node = 0, bsum = 0;
    for (int i = 0; i < kNodalWaves; ++i) {
      node += s_node[i];
      bsum += s_blk[i];
    }
    const int j = (int)(blockIdx.x % kAqSlots);
    atomicAdd(&a.acc[2 * j], node);
    if (bsum) atomicAdd(&a.acc[2 * j + 1], bsum);
  }
}

// Profiling only: 16 bytes per non-zero of M (its entry read from the row
// and from the column structure of graph_parallel.hpp:761-988), spread over
// acc[0 .. kAqSlots).
__global__ __launch_bounds__(256) void k_nnz_bytes(const Slot* __restrict__ kv, int64_t P,
                                                   unsigned long long* acc) {
  long c = 0;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < P;
       j += (int64_t)gridDim.x * blockDim.x)
    c += kv[j].o > 0 ? 1 : 0;
  c = wave_sum(c);
  if (lane_id() == 0 && c)
    atomicAdd(&acc[(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) % kAqSlots],
              16ull * (unsigned long long)c);
}

// M build when every block holds one node (the identity start, B == N): block
// row labels[n] is node n's own list -- out-edges as o, in-edges as i -- so
// one wavefront builds it alone from its neighbour-block table, with no
// contention from other wavefronts (graph_parallel.hpp:703-759).
template <bool kBig>
__device__ __forceinline__ void build_single_one(const GraphView& g, const int32_t* labels,
                                                 const MView& M, const int64_t* big_off,
                                                 int32_t* big_tab, int32_t n, int32_t* tab,
                                                 int32_t* flags) {
  const int lane = lane_id();
  const int64_t o0 = g.out_ptr[n], i0 = g.in_ptr[n];
  const int64_t kco = g.out_ptr[n + 1] - o0, kci = g.in_ptr[n + 1] - i0;
  const int64_t kc = kco + kci;
  const int32_t a = labels[n];
  const Table T = tab_open<kBig>(tab, big_off, big_tab, n, kc);
  const uint32_t mask = (uint32_t)T.cap - 1;
  for (int64_t j = lane; j < kc; j += kWave) {
    int32_t v, w;
    list_entry(g, o0, kco, i0, kc, j, v, w);
    tab_add(T.tk, j < kco ? T.two : T.twi, mask, labels[v], w);
  }
  wave_fence();
  long x0 = 0, x1 = 0, x2 = 0, x3 = 0;
  const int32_t nent = tab_compact(T, -2, -2, x0, x1, x2, x3);
  for (int32_t j = lane; j < nent; j += kWave) m_add<false>(M, a, T.tk[j], T.two[j], T.twi[j], flags);
}

__global__ __launch_bounds__(kWave* kNodalWaves) void k_build_single(GraphView g,
                                                                    const int32_t* labels, MView M,
                                                                    const int64_t* big_off,
                                                                    int32_t* big_tab,
                                                                    int32_t* flags) {
  __shared__ int32_t s_tab[kNodalWaves][3 * kTab];
  const int w = threadIdx.x / kWave;
  const int64_t n = (int64_t)blockIdx.x * kNodalWaves + w;
  if (n >= g.N) return;
  const int64_t kc = (g.out_ptr[n + 1] - g.out_ptr[n]) + (g.in_ptr[n + 1] - g.in_ptr[n]);
  if (kc > kTabMaxDeg)
    build_single_one<true>(g, labels, M, big_off, big_tab, (int32_t)n, s_tab[w], flags);
  else
    build_single_one<false>(g, labels, M, big_off, big_tab, (int32_t)n, s_tab[w], flags);
}

// ---------------------------------------------------------------------------
// Apply accepted moves: block rows of M, degrees, block sizes, batch sum
// (graph_parallel.hpp:433-443; equal to the rebuild from the new labels)
// ---------------------------------------------------------------------------

struct ApplyArgs {
  GraphView g;
  MView M;
  int32_t* labels;
  int32_t* d_out;
  int32_t* d_in;
  int32_t* bsize;
  int32_t beg, end;
  const int64_t* big_off;
  int32_t* big_tab;
  const int32_t* prop_s;
  const long* prop_hi;
  const unsigned long* prop_lo;
  const int32_t* prop_inf;
  const int32_t* prop_acc;
  const int32_t* prop_r;
  const BlockPart* part;  // k_nodal's per-workgroup shares of the batch
  uint32_t nparts;
  Status* status;
  Sweep* sweep;
  unsigned long long* tstamp;  // as NodalArgs::tstamp, for this kernel
  int32_t tstride;
};

__device__ double fx_to_double_dev(i128 v) {
  const bool neg = v < 0;
  const u128 m = neg ? (u128)0 - (u128)v : (u128)v;
  const unsigned long long hi = (unsigned long long)(m >> 64), lo = (unsigned long long)m;
  const double d = ldexp((double)hi, 64 - usap::kLambdaFrac) + ldexp((double)lo, -usap::kLambdaFrac);
  return neg ? -d : d;
}

// A sweep's starting S from k_entropy's per-workgroup partials, stored in
// the device sweep state: the exact int128 sum, then model + data in the same
// IEEE double operations as usap_entropy on the host.
__global__ __launch_bounds__(256) void k_entropy_to_sweep(const long* __restrict__ part,
                                                          int32_t nparts, double model, Sweep* sw) {
  __shared__ long s_hi[4];
  __shared__ unsigned long s_lo[4];
  i128 t = 0;
  for (int32_t i = threadIdx.x; i < nparts; i += blockDim.x)
    t += (i128)(((u128)(unsigned long)part[2 * i] << 64) | (u128)(unsigned long)part[2 * i + 1]);
  t = wave_sum128(t);
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) {
    s_hi[w] = (long)(t >> 64);
    s_lo[w] = (unsigned long)(u128)t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    i128 tot = 0;
    for (int i = 0; i < (int)(blockDim.x / kWave); ++i)
      tot += (i128)(((u128)(unsigned long)s_hi[i] << 64) | (u128)s_lo[i]);
    sw->S = model + fx_to_double_dev(-tot);
  }
}

// graph_parallel.hpp:433-458 after the moves of one batch (sum: exact sum of
// the batch's accepted finite dS, ninf: accepted -inf ones): cumulative dS of
// the sweep, S += dS_sum, the moving-average convergence test.  Same IEEE
// double operations as the host driver.  The batch's error and rebuild flags
// are k_grow's to act on (they may still be raised while this runs).
__device__ void sweep_tail(Sweep* sw, u128 sum, int32_t ninf, int32_t moves,
                           unsigned long long nbytes, unsigned long long abytes) {
  sw->nodal_bytes += nbytes;
  sw->apply_bytes += abytes;
  const u128 cum = (((u128)sw->cum_hi << 64) | sw->cum_lo) + sum;
  sw->cum_lo = (unsigned long long)cum;
  sw->cum_hi = (unsigned long long)(cum >> 64);
  if (ninf > 0) sw->cum_inf = 1;
  const double dS_sum = sw->cum_inf ? -INFINITY : fx_to_double_dev((i128)cum);
  sw->dS_last = dS_sum;
  sw->S += dS_sum;
  sw->done += 1;
  sw->moves += (unsigned long long)moves;
  if (sw->check && -sw->mean < sw->thr * sw->S) atomicExch(&sw->stop, 1);
}

// Moves accepted node n from block r to s: its edges' counts leave row and
// column r for row and column s.  Edges to neighbours that stay put are first
// summed per neighbour block b in the node's table (wo = edges n -> b, wi =
// edges b -> n), so each changed pair of counts gets one 64-bit atomic per
// node (apply_entry); edges to neighbours that move in the same batch are
// applied here, one by one.  Also moves the node's degrees and writes its
// label.  Old labels of nodes inside the batch come from prop_r, so writing
// labels cannot race with their reads.  Returns the table, compacted to its
// nent entries (nent = 0 when none is left).
template <bool kBig>
__device__ __forceinline__ Table apply_table(const ApplyArgs& a, int32_t n, int32_t r, int32_t s,
                                             int64_t o0, int64_t kco, int64_t i0, int64_t kci,
                                             int32_t* tab, int32_t& nent) {
  const int lane = lane_id();
  const GraphView& g = a.g;
  int32_t* flags = &a.status->flags;
  const int64_t kc = kco + kci;
  const Table T = tab_open<kBig>(tab, a.big_off, a.big_tab, n, kc);
  const uint32_t mask = (uint32_t)T.cap - 1;
  long wo_sum = 0, wi_sum = 0;
  for (int64_t base = 0; base < kc; base += kEdgeRegs * kWave) {
    int32_t v[kEdgeRegs], ww[kEdgeRegs], bo[kEdgeRegs], bn[kEdgeRegs];
#pragma unroll
    for (int q = 0; q < kEdgeRegs; ++q)
      list_entry(g, o0, kco, i0, kc, base + q * kWave + lane, v[q], ww[q]);
#pragma unroll
    for (int q = 0; q < kEdgeRegs; ++q) {
      bo[q] = bn[q] = -1;
      if (v[q] < 0) continue;
      if (v[q] >= a.beg && v[q] < a.end) {
        const int32_t x = v[q] - a.beg;
        bo[q] = a.prop_r[x];
        bn[q] = a.prop_acc[x] ? a.prop_s[x] : -2;
      } else {
        bo[q] = a.labels[v[q]];
        bn[q] = -2;
      }
    }
#pragma unroll
    for (int q = 0; q < kEdgeRegs; ++q) {
      if (v[q] < 0) continue;
      const int32_t w = ww[q];
      if (base + q * kWave + lane < kco) {
        wo_sum += w;
        if (bn[q] == -2) {
          tab_add(T.tk, T.two, mask, bo[q], w);
        } else {  // the neighbour moved in this batch too: bo -> bn
          // four independent updates: their rows, then their first keys,
          // are read together before the adds
          const int32_t ra[4] = {r, bo[q], s, bn[q]}, kb[4] = {bo[q], r, bn[q], s};
          const int32_t dO[4] = {-w, 0, w, 0}, dI[4] = {0, -w, 0, w};
          Row R4[4];
          int64_t sl[4];
          int32_t k4[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) R4[i] = m_row(a.M, ra[i]);
#pragma unroll
          for (int i = 0; i < 4; ++i) k4[i] = m_first_key(a.M, R4[i], kb[i], sl[i]);
#pragma unroll
          for (int i = 0; i < 4; ++i)
            m_add_at<true>(a.M, ra[i], R4[i], kb[i], dO[i], dI[i], sl[i], k4[i], flags);
        }
      } else {
        wi_sum += w;
        if (bn[q] == -2) tab_add(T.tk, T.twi, mask, bo[q], w);
        // else: moved as the neighbour's out-edge
      }
    }
  }
  wave_fence();
  long x0 = 0, x1 = 0, x2 = 0, x3 = 0;
  nent = tab_compact(T, -2, -2, x0, x1, x2, x3);
  wo_sum = wave_sum(wo_sum);
  wi_sum = wave_sum(wi_sum);
  if (lane == 0) {
    if (wo_sum) {
      atomicSub(&a.d_out[r], (int32_t)wo_sum);
      atomicAdd(&a.d_out[s], (int32_t)wo_sum);
    }
    if (wi_sum) {
      atomicSub(&a.d_in[r], (int32_t)wi_sum);
      atomicAdd(&a.d_in[s], (int32_t)wi_sum);
    }
    atomicSub(&a.bsize[r], 1);
    atomicAdd(&a.bsize[s], 1);
    a.labels[n] = s;
  }
  return T;
}

// One entry (b, wo, wi) of a moving node's table, node moving r -> s: block
// row r at key b (-wo, -wi), row s at b (+wo, +wi), row b at r (-wi, -wo)
// and at s (+wi, +wo); the four first keys are read together before the adds.
__device__ __forceinline__ void apply_entry(const ApplyArgs& a, int32_t r, Row rM, int32_t s,
                                            Row sM, int32_t b, int32_t wo, int32_t wi) {
  int32_t* flags = &a.status->flags;
  const Row bM = m_row(a.M, b);
  int64_t sl[4];
  int32_t k[4];
  k[0] = m_first_key(a.M, rM, b, sl[0]);  // (M[r][b], M[b][r]) -= (wo, wi)
  k[1] = m_first_key(a.M, sM, b, sl[1]);  // (M[s][b], M[b][s]) += (wo, wi)
  k[2] = m_first_key(a.M, bM, r, sl[2]);  // (M[b][r], M[r][b]) -= (wi, wo)
  k[3] = m_first_key(a.M, bM, s, sl[3]);  // (M[b][s], M[s][b]) += (wi, wo)
  m_add_at<true>(a.M, r, rM, b, -wo, -wi, sl[0], k[0], flags);
  m_add_at<true>(a.M, s, sM, b, wo, wi, sl[1], k[1], flags);
  m_add_at<true>(a.M, b, bM, r, -wi, -wo, sl[2], k[2], flags);
  m_add_at<true>(a.M, b, bM, s, wi, wo, sl[3], k[3], flags);
}

// The batch's sums over all workgroups' shares (nparts of them), by one
// block, then the sweep step.
__device__ __forceinline__ void batch_tail(const BlockPart* part, uint32_t nparts, Sweep* sw) {
  __shared__ unsigned long s_lo[kNodalWaves];
  __shared__ long s_hi[kNodalWaves];
  __shared__ unsigned long long s_nb[kNodalWaves], s_ab[kNodalWaves];
  __shared__ long s_mv[kNodalWaves], s_ni[kNodalWaves];
  const int w = threadIdx.x / kWave;
  i128 sum = 0;
  long nb = 0, ab = 0, mv = 0, ni = 0;
  for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) {
    const BlockPart P = part[i];
    sum += (i128)(((u128)(unsigned long)P.sum_hi << 64) | (u128)P.sum_lo);
    nb += (long)P.nodal_bytes;
    ab += (long)P.apply_bytes;
    mv += P.moves;
    ni += P.ninf;
  }
  sum = wave_sum128(sum);
  nb = wave_sum(nb);
  ab = wave_sum(ab);
  mv = wave_sum(mv);
  ni = wave_sum(ni);
  if (lane_id() == 0) {
    s_lo[w] = (unsigned long)(u128)sum;
    s_hi[w] = (long)(sum >> 64);
    s_nb[w] = (unsigned long long)nb;
    s_ab[w] = (unsigned long long)ab;
    s_mv[w] = mv;
    s_ni[w] = ni;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    u128 tot = 0;
    unsigned long long tnb = 0, tab = 0;
    long tmv = 0, tni = 0;
    for (int i = 0; i < (int)(blockDim.x / kWave); ++i) {
      tot += ((u128)(unsigned long)s_hi[i] << 64) | (u128)s_lo[i];
      tnb += s_nb[i];
      tab += s_ab[i];
      tmv += s_mv[i];
      tni += s_ni[i];
    }
    sweep_tail(sw, tot, (int32_t)tni, (int32_t)tmv, tnb, tab);
  }
  __syncthreads();
}

// Block 0 runs the batch's update step from k_nodal's shares while blocks
// 1.. apply the moves of items (block - 1) * kNodalWaves + wave.
// Block 0 runs the batch's update step; every other block applies the moves
// of its four batch nodes: each moving node's wavefront builds its table,
// then the entries of all the block's tables are spread over all its
// threads (a block holds 0-4 moving nodes, most often 0 or 1, so a node's
// entries no longer wait for its own 64 lanes only).
__global__ __launch_bounds__(kWave* kNodalWaves) void k_apply(ApplyArgs a) {
  __shared__ int32_t s_tab[kNodalWaves][3 * kTab];
  __shared__ int32_t* s_tk[kNodalWaves];  // each moving node's compacted table (LDS or global)
  __shared__ int32_t s_stride[kNodalWaves];
  __shared__ int32_t s_n[kNodalWaves], s_r[kNodalWaves], s_s[kNodalWaves];
  if (blockIdx.x == 0) {
    if (a.tstamp && threadIdx.x == 0) a.tstamp[0] = wall_clock64();
    if (a.sweep->stop_n == 0) batch_tail(a.part, a.nparts, a.sweep);
    if (a.tstamp) {
      __syncthreads();
      if (threadIdx.x == 0) a.tstamp[a.tstride] = wall_clock64();
    }
    return;
  }
  const int w = threadIdx.x / kWave;
  const int64_t item = (int64_t)xcd_block(blockIdx.x - 1, gridDim.x - 1) * kNodalWaves + w;
  const int64_t n64 = (int64_t)a.beg + item;
  const bool in = n64 < a.end;
  const int32_t n = in ? (int32_t)n64 : a.beg;
  // whether the batch runs, the proposal and the node's list bounds in one
  // round of loads
  const int32_t stop = a.sweep->stop_n;  // the same for every block of this launch
  const int32_t acc = in ? a.prop_acc[item] : 0;
  const int32_t pr = in ? a.prop_r[item] : 0, ps = in ? a.prop_s[item] : 0;
  const int64_t o0 = a.g.out_ptr[n], i0 = a.g.in_ptr[n];
  const int64_t kco = a.g.out_ptr[n + 1] - o0, kci = a.g.in_ptr[n + 1] - i0;
  if (stop) return;
  if (a.tstamp && threadIdx.x == 0) a.tstamp[blockIdx.x] = wall_clock64();
  int32_t nent = 0;
  Table T{s_tab[w], s_tab[w] + kTab, s_tab[w] + 2 * kTab, kTab};
  if (in && acc) {
    if (kco + kci > kTabMaxDeg)
      T = apply_table<true>(a, n, pr, ps, o0, kco, i0, kci, s_tab[w], nent);
    else
      T = apply_table<false>(a, n, pr, ps, o0, kco, i0, kci, s_tab[w], nent);
  }
  if (lane_id() == 0) {
    s_n[w] = nent;
    s_r[w] = pr;
    s_s[w] = ps;
    s_tk[w] = T.tk;
    s_stride[w] = (int32_t)(T.two - T.tk);
  }
  __syncthreads();
  int32_t off[kNodalWaves + 1];
  off[0] = 0;
#pragma unroll
  for (int i = 0; i < kNodalWaves; ++i) off[i + 1] = off[i] + s_n[i];
  const int32_t total = off[kNodalWaves];
  for (int32_t e = threadIdx.x; e < total; e += blockDim.x) {
    int wv = 0;
#pragma unroll
    for (int i = 1; i < kNodalWaves; ++i) wv += e >= off[i];
    const int32_t j = e - off[wv];
    const int32_t* tk = s_tk[wv];
    const int32_t st = s_stride[wv];
    const int32_t b = tk[j], wo = tk[st + j], wi = tk[2 * st + j];
    if (wo == 0 && wi == 0) continue;
    const int32_t r = s_r[wv], s = s_s[wv];
    apply_entry(a, r, m_row(a.M, r), s, m_row(a.M, s), b, wo, wi);
  }
  if (a.tstamp) {
    __syncthreads();
    if (threadIdx.x == 0) a.tstamp[a.tstride + blockIdx.x] = wall_clock64();
  }
}

// Rows whose key count passed 1/slack of the slab or that have parked
// increments (each of the 512 workgroups scans its slice of the rows),
// handled between two batches, dead keys (both counts 0) dropped: a
// row whose live keys and parked increments fit 1/(2 slack) of its slab (and
// at most kCompact slots) is compacted in place through LDS; any other row
// moves to a slab >= 2*slack*live+16 (direct once that reaches B/8) at the
// free tail of the pool, and the row's parked increments are replayed into it.  A
// new slab of at most kGrowLds slots is laid out in LDS (its slots claimed
// meanwhile) and written out whole; the old one is left as zeros.  Bucket
// sums are unchanged.  Block 0 also acts on the batch's flags.
constexpr int kCompact = 4096;
constexpr int kGrowLds = 2048;    // slabs up to this many slots are laid out in LDS
constexpr int kGrowBlocks = 512;  // k_grow's grid
constexpr int kGrowScan = 1024;   // rows a k_grow workgroup checks per round (4 per thread)
constexpr int32_t kOvfCap = 1 << 22;  // parked increments

struct GrowSide {
  int64_t* off;
  int32_t* cap;
  int32_t* nkeys;
  Slot* kv;
  int32_t* gpend;
  const int32_t* ovf;
  int32_t ovf_cap;
  GrowState* gst;
  int64_t P;
  int32_t slack;
};

struct GrowArgs {
  GrowSide side;
  int32_t B;
  Sweep* sweep;
  Status* status;
  unsigned long long* tstamp;  // as NodalArgs::tstamp, for this kernel
  int32_t tstride;
};

// Replays the parked increments of row a into its slab (o, c) (the only block
// touching row a in this launch); an increment that still finds no slot asks
// the host for a rebuild.
__device__ void replay_row(const GrowSide& g, int32_t a, int64_t o, int32_t c, int32_t B,
                           Sweep* sw) {
  const int32_t no = min(g.gst->ovf_n, g.ovf_cap);
  for (int32_t i = threadIdx.x; i < no; i += blockDim.x) {
    if (g.ovf[i] != a) continue;
    const int32_t b = g.ovf[g.ovf_cap + i];
    const int32_t dout = g.ovf[2 * g.ovf_cap + i], din = g.ovf[3 * g.ovf_cap + i];
    if (c == B) {
      cnt_add(&g.kv[o + b], dout, din);
      continue;
    }
    const uint32_t mask = (uint32_t)c - 1;
    uint32_t h = hash32(b) & mask;
    bool done = false;
    for (int32_t probe = 0; probe < c && !done; ++probe) {
      int32_t k = __hip_atomic_load(&g.kv[o + h].k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (k < 0) {
        k = atomicCAS(&g.kv[o + h].k, -1, b);
        if (k == -1) {
          atomicAdd(&g.nkeys[a], 1);
          k = b;
        }
      }
      if (k == b) {
        cnt_add(&g.kv[o + h], dout, din);
        done = true;
      }
      h = (h + 1) & mask;
    }
    if (!done) {  // no room (the pool ran out): rebuild M
      atomicOr(&sw->flags, kFlagFull);
      atomicCAS(&sw->stop, 0, 2);
    }
  }
}

// The live entries s_k/s_o/s_i[0, live) laid out as a slab of nc slots in
// LDS t_k/t_o/t_i (hashed, or by key when direct), for one coalesced
// write-out: no per-entry CAS round trip to the new slab.
__device__ void lds_slab(const int32_t* s_k, const int32_t* s_o, const int32_t* s_i,
                         int32_t live, int32_t* t_k, int32_t* t_o, int32_t* t_i, int32_t nc,
                         bool direct) {
  for (int32_t j = threadIdx.x; j < nc; j += blockDim.x) {
    t_k[j] = -1;
    t_o[j] = 0;
    t_i[j] = 0;
  }
  __syncthreads();
  const uint32_t mask = (uint32_t)nc - 1;
  for (int32_t i = threadIdx.x; i < live; i += blockDim.x) {
    const int32_t k = s_k[i];
    uint32_t h = (uint32_t)k;
    if (!direct) {
      h = hash32(k) & mask;
      while (atomicCAS(&t_k[h], -1, k) != -1) h = (h + 1) & mask;
    }
    t_o[h] = s_o[i];
    t_i[h] = s_i[i];
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_grow(GrowArgs ga) {
  __shared__ int32_t s_k[kCompact];  // the row's live entries
  __shared__ int32_t s_o[kCompact];
  __shared__ int32_t s_i[kCompact];
  __shared__ int32_t t_k[kGrowLds];  // its new slab, when it fits
  __shared__ int32_t t_o[kGrowLds];
  __shared__ int32_t t_i[kGrowLds];
  __shared__ int32_t s_q[kGrowScan];  // rows of the scanned chunk to handle
  __shared__ long long s_base;
  __shared__ int32_t s_cap, s_live, s_nq;
  const int32_t B = ga.B;
  const GrowSide& g = ga.side;
  if (ga.tstamp && threadIdx.x == 0) ga.tstamp[blockIdx.x] = wall_clock64();
  // the batch's flags (raised by its k_nodal and k_apply, both complete):
  // errors override convergence, which overrides a rebuild request, as the
  // decision comes after the batch
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int32_t f = ga.status->flags;
    if (f) {
      atomicOr(&ga.sweep->flags, f);
      if (f & (kFlagMissing | kFlagNegative))
        atomicExch(&ga.sweep->stop, 3);
      else if (f & kFlagFull)
        atomicCAS(&ga.sweep->stop, 0, 2);
      ga.status->flags = 0;
    }
  }
  // the rows to grow or compact, found by a scan of this workgroup's slice
  // of the rows: hashed rows past 1/slack load, or with parked increments
  // (k_apply only counts keys and parked increments; it queues nothing)
  const int32_t per = (B + (int32_t)gridDim.x - 1) / (int32_t)gridDim.x;
  const int32_t lo = (int32_t)blockIdx.x * per, hi = min(B, lo + per);
  for (int32_t base = lo; base < hi; base += kGrowScan) {
    if (threadIdx.x == 0) s_nq = 0;
    __syncthreads();
    int32_t cq[kGrowScan / 256], nq_k[kGrowScan / 256], pq[kGrowScan / 256];
#pragma unroll
    for (int q = 0; q < kGrowScan / 256; ++q) {
      const int32_t r = base + q * 256 + (int32_t)threadIdx.x;
      cq[q] = B;
      nq_k[q] = 0;
      pq[q] = 0;
      if (r < hi) {
        cq[q] = g.cap[r];
        nq_k[q] = g.nkeys[r];
        pq[q] = g.gpend[r];
      }
    }
#pragma unroll
    for (int q = 0; q < kGrowScan / 256; ++q)
      if (cq[q] < B && ((int64_t)g.slack * nq_k[q] > cq[q] || pq[q] > 0))
        s_q[atomicAdd(&s_nq, 1)] = base + q * 256 + (int32_t)threadIdx.x;
    __syncthreads();
    const int32_t nq = s_nq;
    for (int32_t qi = 0; qi < nq; ++qi) {
    const int32_t a = s_q[qi];
    const int64_t o = g.off[a];
    const int32_t c = g.cap[a];
    const int32_t pend = g.gpend[a];  // parked increments (at most one new key each)
    const bool dold = c == B;
    // a hashed slab of at most kCompact slots has its live entries gathered
    // into LDS by the counting pass
    const bool gather = !dold && c <= kCompact;
    if (threadIdx.x == 0) s_live = 0;
    __syncthreads();
    for (int32_t j = threadIdx.x; j < c; j += blockDim.x) {
      const Slot x = g.kv[o + j];
      const int32_t k = dold ? j : x.k;
      if (k >= 0 && (x.o != 0 || x.i != 0)) {
        const int32_t i = atomicAdd(&s_live, 1);
        if (gather) {
          s_k[i] = k;
          s_o[i] = x.o;
          s_i[i] = x.i;
        }
      }
    }
    __syncthreads();
    const int32_t live = s_live;
    int64_t o2 = o;  // the row's slab after this step
    int32_t c2 = c;
    if (gather && 2 * g.slack * (live + pend) <= c) {
      // in place: the slab rewritten with the live entries only
      if (c <= kGrowLds) {
        lds_slab(s_k, s_o, s_i, live, t_k, t_o, t_i, c, false);
        for (int32_t j = threadIdx.x; j < c; j += blockDim.x)
          g.kv[o + j] = Slot{t_k[j], 0, t_o[j], t_i[j]};
      } else {
        for (int32_t j = threadIdx.x; j < c; j += blockDim.x) g.kv[o + j] = Slot{-1, 0, 0, 0};
        __threadfence_block();
        __syncthreads();
        const uint32_t mask = (uint32_t)c - 1;
        for (int32_t i = threadIdx.x; i < live; i += blockDim.x) {
          uint32_t h = hash32(s_k[i]) & mask;
          while (atomicCAS(&g.kv[o + h].k, -1, s_k[i]) != -1) h = (h + 1) & mask;
          g.kv[o + h].o = s_o[i];
          g.kv[o + h].i = s_i[i];
        }
      }
      __threadfence_block();
      __syncthreads();
      if (threadIdx.x == 0) {
        g.nkeys[a] = live;
        g.gst->grows += 1;
        g.gst->compactions += 1;
      }
    } else {
      int64_t nc = 32;
      while (nc < 2 * g.slack * ((int64_t)live + pend) + 16) nc <<= 1;
      if (nc < 2 * (int64_t)c) nc = 2 * (int64_t)c;
      if (8 * nc >= B) nc = B;
      // the slots are claimed while the new slab is laid out in LDS; a
      // direct row never moves
      unsigned long long base = 0;
      if (threadIdx.x == 0 && !dold) base = atomicAdd(&g.gst->used, (unsigned long long)nc);
      const bool inlds = gather && nc <= kGrowLds;
      if (inlds) lds_slab(s_k, s_o, s_i, live, t_k, t_o, t_i, (int32_t)nc, nc == B);
      if (threadIdx.x == 0) {
        s_cap = 0;
        if (!dold) {
          if ((int64_t)base + nc > g.P) {
            atomicAdd(&g.gst->used, (unsigned long long)(-nc));  // give the slots back
            g.gst->nospace = 1;
            g.gst->nospace_events += 1;
            atomicOr(&ga.sweep->flags, kFlagFull);
            atomicCAS(&ga.sweep->stop, 0, 2);
          } else {
            s_cap = (int32_t)nc;
            s_base = (long long)base;
          }
        }
      }
      __syncthreads();
      const int32_t ncap = s_cap;
      if (ncap > 0) {
        const int64_t nb = s_base;
        if (inlds) {
          for (int32_t j = threadIdx.x; j < ncap; j += blockDim.x)
            g.kv[nb + j] = Slot{t_k[j], 0, t_o[j], t_i[j]};
          // the old slab stays in the pool as zeros until a rebuild
          for (int32_t j = threadIdx.x; j < c; j += blockDim.x) g.kv[o + j] = Slot{-1, 0, 0, 0};
        } else {
          for (int32_t i = threadIdx.x; i < ncap; i += blockDim.x) g.kv[nb + i] = Slot{-1, 0, 0, 0};
          __threadfence_block();
          __syncthreads();
          const bool dnew = ncap == B;
          const uint32_t mask = (uint32_t)ncap - 1;
          for (int32_t j = threadIdx.x; j < c; j += blockDim.x) {
            const Slot x = g.kv[o + j];
            if (x.k < 0 || (x.o == 0 && x.i == 0)) continue;
            g.kv[o + j].o = 0;
            g.kv[o + j].i = 0;
            if (dnew) {
              g.kv[nb + x.k].o = x.o;
              g.kv[nb + x.k].i = x.i;
              continue;
            }
            uint32_t h = hash32(x.k) & mask;
            while (atomicCAS(&g.kv[nb + h].k, -1, x.k) != -1) h = (h + 1) & mask;
            g.kv[nb + h].o = x.o;
            g.kv[nb + h].i = x.i;
          }
        }
        __threadfence_block();
        __syncthreads();
        if (threadIdx.x == 0) {
          g.off[a] = nb;
          g.cap[a] = ncap;
          g.nkeys[a] = live;
          g.gst->grows += 1;
        }
        o2 = nb;
        c2 = ncap;
      }
    }
    if (pend > 0) replay_row(g, a, o2, c2, B, ga.sweep);
    __syncthreads();
    if (threadIdx.x == 0) g.gpend[a] = 0;
    __syncthreads();
    }
  }
  if (ga.tstamp) {
    __syncthreads();
    if (threadIdx.x == 0) ga.tstamp[ga.tstride + blockIdx.x] = wall_clock64();
  }
}

}  // namespace

// ===========================================================================
// Host side
// ===========================================================================

struct DevSpM {
  int64_t* off = nullptr;
  int32_t* cap = nullptr;
  int32_t* nkeys = nullptr;
  Slot* kv = nullptr;     // row slabs [P]
  int32_t* bs = nullptr;  // [bs_rows][kNBK]
  int32_t bs_rows = 0;
  int32_t* gpend = nullptr;  // [N]
  int32_t* ovf = nullptr;    // [4][ovf_cap]
  int32_t ovf_cap = 0;       // parked increments that fit
  GrowState* gst = nullptr;
  int64_t P = 0;      // allocated slots
  int64_t total = 0;  // slots in use after the last rebuild
  int32_t slack = 2;  // usap_params.slab_slack
};

struct usap_engine {
  usap_params p{};
  hipStream_t stream = nullptr;
  bool own_stream = true;      // false: a worker context's stream, kept at destroy
  hipMemPool_t pool = nullptr;  // a worker context's pool, else the device's default pool
  std::string err;
  // host copy of the graph
  int32_t N = 0;
  int64_t E = 0;
  bool weighted = false;
  std::vector<int64_t> h_out_ptr, h_in_ptr;
  std::vector<int32_t> h_out_idx, h_in_idx, h_out_w, h_in_w;
  usap_graph hg{};
  bool host_stale = false;  // lists appended on the device since the host copy was made
  int64_t Wtot = 0;
  int64_t lam_cap = 0;  // entries of the device lambda table
  // device graph
  int64_t *out_ptr = nullptr, *in_ptr = nullptr;
  int32_t *out_idx = nullptr, *in_idx = nullptr, *out_w = nullptr, *in_w = nullptr;
  int32_t* out_src = nullptr;
  int64_t* lam = nullptr;
  // partition state
  int32_t B = 0;
  bool has_labels = false;
  int32_t* labels = nullptr;
  int32_t *d_out = nullptr, *d_in = nullptr, *ecnt_out = nullptr, *ecnt_in = nullptr,
          *bsize = nullptr;
  // nodes grouped by block for the per-row build: offsets, arrival cursors, nodes
  int32_t *bl_off = nullptr, *bl_cur = nullptr, *bl_nodes = nullptr;
  DevSpM M;  // block rows (both orientations of M)
  int64_t* tmpM = nullptr;
  void* scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
  // merge
  int32_t* best_s = nullptr;
  long* best_hi = nullptr;
  unsigned long* best_lo = nullptr;
  int32_t merge_B = -1;  // B of the last merge_propose (-1: none)
  int32_t* map = nullptr;
  unsigned long long *sort_key[2] = {nullptr, nullptr};  // carry-out sort buffers [N]
  int32_t* sort_idx[2] = {nullptr, nullptr};
  void* sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  // nodal
  int32_t* prop_s = nullptr;
  long* prop_hi = nullptr;
  unsigned long* prop_lo = nullptr;
  int32_t* prop_inf = nullptr;
  int32_t* prop_acc = nullptr;
  int32_t last_beg = 0, last_end = 0;
  // pre-drawn uniform stream (usap_set_uniforms): 4 words per draw, or null
  uint64_t* uni = nullptr;
  int64_t uni_n = 0;    // draws held
  int64_t uni_pos = 0;  // draws consumed
  Status* status = nullptr;   // the running batch
  BlockPart* part = nullptr;  // per-workgroup shares of the running batch
  Sweep* sweep = nullptr;     // device sweep state
  Sweep* h_sweep = nullptr;   // pinned host copy
  Sweep* h_chunk[2] = {nullptr, nullptr};  // sweep state after alternate launch chunks
  hipEvent_t chunk_ev[2] = {nullptr, nullptr};
  int32_t* prop_r = nullptr;  // old label of each batch node
  // reference-algorithm byte counts (profiling): drawn neighbour's block per
  // batch node, per-block batch stamps, device totals [nodal per-node,
  // nodal per-block, merge non-zeros]
  int32_t* prop_u = nullptr;
  int32_t* aq_mark = nullptr;
  unsigned long long* aq_acc = nullptr;
  int32_t aq_stamp = 0;
  unsigned long long* tstamp = nullptr;  // profiled batch launches: workgroup stamps
  size_t tstamp_n = 0;                   // its length
  int32_t tstamp_G = 0;                  // k_nodal / k_apply workgroups per slot this round
  // captured sweep (usap_params.graphs): replayed while its key matches
  hipGraphExec_t gexec = nullptr;
  int32_t g_B = -1, g_batch = -1, g_nb = -1;
  uint64_t g_epoch = 0;
  uint64_t epoch = 0;  // bumped by every device allocation (pointers may change)
  int64_t graph_captures = 0;
  int32_t* build_flags = nullptr;  // [0] build overflow flags, [1] blocks holding != 1 node, [2] distinct block pairs
  uint32_t* pair_key[2] = {nullptr, nullptr};  // block-pair builds [E]
  int32_t* pair_val[2] = {nullptr, nullptr};
  int64_t pair_cap = 0;  // edges the pair buffers hold
  void* pair_tmp = nullptr;
  size_t pair_tmp_bytes = 0;
  int64_t* big_off = nullptr;
  int32_t* big_tab = nullptr;
  // entropy
  long* ent_part = nullptr;
  long* h_ent_part = nullptr;
  // snapshots
  int32_t* snap_labels[kNumSnap] = {};
  int32_t snap_B[kNumSnap] = {};
  double snap_S[kNumSnap] = {};
  int32_t rebuilds = 0;
  // profiling (HIP events on the engine's stream)
  bool prof = false;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  bool all_direct = false;  // every row of M direct-mapped (set by each rebuild)
  usap_kernel_prof kp[USAP_NKERNELS] = {};
  std::vector<usap_phase> phase_log;
};

namespace {

int set_err(usap_engine* e, int code, const std::string& msg) {
  if (e) e->err = msg;
  return code;
}

#define HIPCHK(expr)                                                                    \
  do {                                                                                  \
    hipError_t _st = (expr);                                                            \
    if (_st != hipSuccess)                                                              \
      return set_err(e, USAP_EHIP, std::string(#expr) + ": " + hipGetErrorString(_st)); \
  } while (0)

// The runtime's resource calls (stream-ordered allocations and frees, pools,
// streams, events, pinned host buffers, graph instantiation) are made under
// one process-wide lock.  Engines on several host threads (batched and served
// graphs) make these calls concurrently; test processes running several
// engines on host threads have ended in crashes inside native threads with
// no GPU fault and no cause found in this library, so the calls that create,
// grow or release runtime state are kept from overlapping.  They are short
// and few per graph; kernel launches and copies are not locked.
std::mutex g_hip_mu;

// Device memory is stream-ordered (hipMallocAsync / hipFreeAsync on the
// engine's stream): creating and destroying engines never synchronises the
// device, so many engines can come and go while others run (batched and
// served graphs).  The default pool keeps freed memory for reuse.
void prepare_pool(int device) {
  static std::once_flag once;
  std::call_once(once, [device]() {
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
      uint64_t keep = UINT64_MAX;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
  });
}

template <typename T>
int dalloc(usap_engine* e, T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  e->epoch++;
  std::unique_lock<std::mutex> lk(g_hip_mu);
  hipError_t st = e->pool ? hipMallocFromPoolAsync((void**)p, n * sizeof(T), e->pool, e->stream)
                          : hipMallocAsync((void**)p, n * sizeof(T), e->stream);
  lk.unlock();
  if (st != hipSuccess) {
    *p = nullptr;
    return set_err(e, USAP_ENOMEM, std::string("hipMallocAsync: ") + hipGetErrorString(st));
  }
  return USAP_OK;
}

// dalloc for untyped scratch (rocPRIM temporary storage).
int dalloc_raw(usap_engine* e, void** p, size_t bytes) {
  *p = nullptr;
  e->epoch++;
  std::unique_lock<std::mutex> lk(g_hip_mu);
  hipError_t st = e->pool ? hipMallocFromPoolAsync(p, bytes ? bytes : 1, e->pool, e->stream)
                          : hipMallocAsync(p, bytes ? bytes : 1, e->stream);
  lk.unlock();
  if (st != hipSuccess) {
    *p = nullptr;
    return set_err(e, USAP_ENOMEM, std::string("hipMallocAsync: ") + hipGetErrorString(st));
  }
  return USAP_OK;
}

template <typename T>
void dfree(usap_engine* e, T*& p) {
  if (p) {
    std::lock_guard<std::mutex> lk(g_hip_mu);
    (void)hipFreeAsync((void*)p, e->stream);
  }
  p = nullptr;
}

// Pinned host buffers are recycled through a process-wide free list instead
// of being freed (hipHostFree may wait for the device).
std::mutex g_pin_mu;
std::vector<std::pair<size_t, void*>> g_pin_free;

void* pin_get(size_t bytes) {
  {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    for (size_t i = 0; i < g_pin_free.size(); ++i)
      if (g_pin_free[i].first == bytes) {
        void* q = g_pin_free[i].second;
        g_pin_free.erase(g_pin_free.begin() + (long)i);
        return q;
      }
  }
  void* q = nullptr;
  std::lock_guard<std::mutex> lk(g_hip_mu);
  if (hipHostMalloc(&q, bytes) != hipSuccess) return nullptr;
  return q;
}

void pin_put(void* q, size_t bytes) {
  if (!q) return;
  std::lock_guard<std::mutex> lk(g_pin_mu);
  g_pin_free.emplace_back(bytes, q);
}

inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

int32_t bucket_shift(int32_t B) {
  int32_t sh = 0;
  while (((B - 1) >> sh) >= kNBK) ++sh;
  return sh;
}

MView view(const DevSpM& m, int32_t B) {
  MView v;
  v.off = m.off;
  v.cap = m.cap;
  v.nkeys = m.nkeys;
  v.kv = m.kv;
  v.bs = m.bs;
  v.gpend = m.gpend;
  v.ovf = m.ovf;
  v.ovf_cap = m.ovf_cap;
  v.gst = m.gst;
  v.B = B;
  v.bshift = bucket_shift(B);
  v.slack = m.slack;
  return v;
}

GraphView gview(const usap_engine* e) {
  GraphView g;
  g.N = e->N;
  g.out_ptr = e->out_ptr;
  g.out_idx = e->out_idx;
  g.out_w = e->out_w;
  g.in_ptr = e->in_ptr;
  g.in_idx = e->in_idx;
  g.in_w = e->in_w;
  return g;
}

double fx_to_double(i128 v) {
  const bool neg = v < 0;
  const u128 m = neg ? (u128)0 - (u128)v : (u128)v;
  const uint64_t hi = (uint64_t)(m >> 64), lo = (uint64_t)m;
  const double d = std::ldexp((double)hi, 64 - usap::kLambdaFrac) +
                   std::ldexp((double)lo, -usap::kLambdaFrac);
  return neg ? -d : d;
}

i128 mk128(long hi, unsigned long lo) { return (i128)(((u128)(unsigned long)hi << 64) | (u128)lo); }

// elapsed ms between two recorded events (0 when profiling is off)
float ev_ms(usap_engine* e, int a, int b) {
  float ms = 0.f;
  if (e->prof && hipEventElapsedTime(&ms, e->ev[a], e->ev[b]) != hipSuccess) ms = 0.f;
  return ms;
}

void ev_rec(usap_engine* e, int i) {
  if (e->prof) (void)hipEventRecord(e->ev[i], e->stream);
}

void prof_add(usap_engine* e, int k, double ms, int64_t launches, double bytes) {
  if (!e->prof) return;
  e->kp[k].ms += ms;
  e->kp[k].launches += launches;
  e->kp[k].bytes += bytes;
}

// The block-pair build (k_pair_keys, sort, sum per pair, k_insert_pairs).
int build_by_pairs(usap_engine* e) {
  const int32_t B = e->B;
  const int64_t E = e->E;
  hipStream_t st = e->stream;
  if (e->pair_cap < E) {  // first use, or the graph grew (usap_set_graph)
    for (int i = 0; i < 2; ++i) {
      dfree(e, e->pair_key[i]);
      dfree(e, e->pair_val[i]);
      if (int rc = dalloc(e, &e->pair_key[i], (size_t)E)) return rc;
      if (int rc = dalloc(e, &e->pair_val[i], (size_t)E)) return rc;
    }
    e->pair_cap = E;
  }
  int bits = 1;
  while (((uint64_t)1 << bits) < (uint64_t)B * (uint64_t)B) ++bits;
  size_t b1 = 0, b2 = 0;
  HIPCHK(rocprim::radix_sort_pairs(nullptr, b1, e->pair_key[0], e->pair_key[1], e->pair_val[0],
                                   e->pair_val[1], (size_t)E, 0, bits, st));
  HIPCHK(rocprim::reduce_by_key(nullptr, b2, e->pair_key[1], e->pair_val[1], (size_t)E,
                                e->pair_key[0], e->pair_val[0], e->build_flags + 2,
                                rocprim::plus<int32_t>(), rocprim::equal_to<uint32_t>(), st));
  const size_t need = std::max(b1, b2);
  if (need > e->pair_tmp_bytes) {
    dfree(e, e->pair_tmp);
    if (int rc = dalloc_raw(e, &e->pair_tmp, need)) return rc;
    e->pair_tmp_bytes = need;
  }
  const unsigned g = std::min<int64_t>(blocks_for(E, 256), 16384);
  k_pair_keys<<<g, 256, 0, st>>>(E, e->out_src, e->out_idx, e->out_w, e->labels, B, e->pair_key[0],
                                 e->pair_val[0]);
  HIPCHK(hipGetLastError());
  size_t bytes = e->pair_tmp_bytes;
  HIPCHK(rocprim::radix_sort_pairs(e->pair_tmp, bytes, e->pair_key[0], e->pair_key[1], e->pair_val[0],
                                   e->pair_val[1], (size_t)E, 0, bits, st));
  bytes = e->pair_tmp_bytes;
  HIPCHK(rocprim::reduce_by_key(e->pair_tmp, bytes, e->pair_key[1], e->pair_val[1], (size_t)E,
                                e->pair_key[0], e->pair_val[0], e->build_flags + 2,
                                rocprim::plus<int32_t>(), rocprim::equal_to<uint32_t>(), st));
  k_insert_pairs<<<g, 256, 0, st>>>(E, e->pair_key[0], e->pair_val[0], e->build_flags + 2,
                                    view(e->M, B), e->build_flags);
  HIPCHK(hipGetLastError());
  return USAP_OK;
}

// Full M build from the labels (graph_parallel.hpp:703-759).
// The M build for B >= kRowBuildB: nodes grouped by block (offsets by an
// exclusive scan of the block sizes, then a scatter), k_build_rows for the
// rows it lays out in LDS, per-edge inserts for the others.
int build_rows(usap_engine* e) {
  const int32_t B = e->B;
  hipStream_t st = e->stream;
  size_t bytes = 0;
  HIPCHK(rocprim::exclusive_scan(nullptr, bytes, e->bsize, e->bl_off, (int32_t)0, (size_t)B,
                                 rocprim::plus<int32_t>(), st));
  if (bytes > e->scan_tmp_bytes) {
    dfree(e, e->scan_tmp);
    if (int rc = dalloc_raw(e, &e->scan_tmp, bytes)) return rc;
    e->scan_tmp_bytes = bytes;
  }
  bytes = e->scan_tmp_bytes;
  HIPCHK(rocprim::exclusive_scan(e->scan_tmp, bytes, e->bsize, e->bl_off, (int32_t)0, (size_t)B,
                                 rocprim::plus<int32_t>(), st));
  k_node_lists<<<std::min<unsigned>(blocks_for(e->N, 256), 4096u), 256, 0, st>>>(
      e->N, e->labels, e->bl_off, e->bl_cur, e->bl_nodes);
  k_build_rows<<<blocks_for(B, kNodalWaves), kWave * kNodalWaves, 0, st>>>(
      gview(e), e->labels, e->bl_off, e->bl_nodes, view(e->M, B), e->build_flags);
  k_insert_edges<true><<<std::min<int64_t>(blocks_for(e->E, 256), 16384), 256, 0, st>>>(
      e->E, e->out_src, e->out_idx, e->out_w, e->labels, view(e->M, B), e->build_flags);
  HIPCHK(hipGetLastError());
  return USAP_OK;
}

int rebuild_M(usap_engine* e) {
  const int32_t B = e->B;
  hipStream_t st = e->stream;
  ev_rec(e, 0);
  {
    ClearArgs c;
    int32_t* arr[kClearArrays] = {e->d_out, e->d_in, e->ecnt_out, e->ecnt_in,
                                  e->bsize, e->M.nkeys, e->M.gpend, e->bl_cur};
    for (int i = 0; i < kClearArrays; ++i) c.p[i] = arr[i];
    c.n = B;
    c.flags = e->build_flags;
    k_clear_build<<<std::max(1u, std::min(blocks_for(B, 256), 1024u)), 256, 0, st>>>(c);
  }
  if (e->N > 0) {
    const unsigned g = std::min<unsigned>(blocks_for(e->N, 256), B <= kDegLds ? 256u : 4096u);
    k_degrees<<<g, 256, 0, st>>>(gview(e), e->labels, B, e->d_out, e->d_in, e->ecnt_out,
                                 e->ecnt_in, e->bsize);
  }
  k_caps<<<blocks_for((int64_t)B + 1, 256), 256, 0, st>>>(B, e->p.slab_slack, e->ecnt_out,
                                                          e->ecnt_in, e->M.cap, e->tmpM);
  HIPCHK(hipGetLastError());
  size_t bytes = 0;
  HIPCHK(rocprim::exclusive_scan(nullptr, bytes, e->tmpM, e->M.off, (int64_t)0, (size_t)B + 1,
                                 rocprim::plus<int64_t>(), st));
  if (bytes > e->scan_tmp_bytes) {
    dfree(e, e->scan_tmp);
    if (int rc = dalloc_raw(e, &e->scan_tmp, bytes)) return rc;
    e->scan_tmp_bytes = bytes;
  }
  bytes = e->scan_tmp_bytes;
  HIPCHK(rocprim::exclusive_scan(e->scan_tmp, bytes, e->tmpM, e->M.off, (int64_t)0, (size_t)B + 1,
                                 rocprim::plus<int64_t>(), st));
  // B == N: is every block a single node (rows built per node)?
  const bool maybe_single = B == e->N && e->E > 0;
  int32_t nonsingle = 1;
  if (maybe_single) {
    k_count_nonsingle<<<std::min<unsigned>(blocks_for(B, 256), 1024u), 256, 0, st>>>(
        B, e->bsize, e->build_flags + 1);
    HIPCHK(hipMemcpyAsync(&nonsingle, e->build_flags + 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  }
  int64_t tot = 0;
  HIPCHK(hipMemcpyAsync(&tot, e->M.off + B, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  DevSpM& m = e->M;
  m.total = tot;
  m.slack = e->p.slab_slack;
  // every row direct-mapped (cap B each; a hashed row's cap is below B/8):
  // no insert probes, so no row ever needs growth or an increment parked until
  // the next rebuild, and k_grow only has the batch's error flags to fold
  e->all_direct = tot == (int64_t)B * B;
  if (3 * m.total > m.P) {  // two thirds of the pool is headroom for rows that grow
    dfree(e, m.kv);
    const int64_t P = 3 * m.total + (int64_t)4 * B + 4096;
    if (int rc = dalloc(e, &m.kv, (size_t)P)) return rc;
    m.P = P;
  }
  if (m.total > 0) k_fill<<<std::min<int64_t>(blocks_for(m.total, 256), 8192), 256, 0, st>>>(
      m.kv, m.total);
  k_reset_grow<<<1, 1, 0, st>>>(m.gst, (unsigned long long)m.total);
  if (m.bs_rows < B) {
    dfree(e, m.bs);
    if (int rc = dalloc(e, &m.bs, (size_t)B * kNBK)) return rc;
    m.bs_rows = B;
  }
  if (e->E > 0) {
    if (maybe_single && nonsingle == 0)
      k_build_single<<<blocks_for(e->N, kNodalWaves), kWave * kNodalWaves, 0, st>>>(
          gview(e), e->labels, view(e->M, B), e->big_off, e->big_tab, e->build_flags);
    else if (B <= kSmallB)
      k_insert_small<<<256, 256, 0, st>>>(e->E, e->out_src, e->out_idx, e->out_w, e->labels,
                                          view(e->M, B));
    else if (B <= kPairSortB) {
      if (int rc = build_by_pairs(e)) return rc;
    } else if (B < kRowBuildB) {
      k_insert_edges<false><<<std::min<int64_t>(blocks_for(e->E, 256), 16384), 256, 0, st>>>(
          e->E, e->out_src, e->out_idx, e->out_w, e->labels, view(e->M, B), e->build_flags);
    } else {
      if (int rc = build_rows(e)) return rc;
    }
  }
  k_bucket_sums<<<blocks_for(B, 4), 256, 0, st>>>(view(e->M, B));
  HIPCHK(hipGetLastError());
  ev_rec(e, 1);
  int32_t flags = 0;
  HIPCHK(hipMemcpyAsync(&flags, e->build_flags, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  prof_add(e, USAP_K_BUILD, ev_ms(e, 0, 1), 1, 0.0);
  if (flags & (kFlagFull | kFlagMissing))
    return set_err(e, USAP_ESTATE, "M build overflowed a row slab");
  e->rebuilds++;
  return USAP_OK;
}

int check_engine(usap_engine* e) {
  if (!e) return USAP_EINVAL;
  if (!e->has_labels) return set_err(e, USAP_ESTATE, "labels not set");
  return USAP_OK;
}

}  // namespace

namespace {
int refresh_host_graph(usap_engine* e);
}  // namespace

namespace usap {
namespace {
std::mutex g_err_mu;
std::string g_err;
thread_local std::string t_err;
}  // namespace

void set_detached_error(const std::string& msg) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  if (g_err.empty()) g_err = msg;  // the first failure is the informative one
}

void clear_detached_error() {
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_err.clear();
}

const char* detached_error() {
  std::lock_guard<std::mutex> lk(g_err_mu);
  t_err = g_err.empty() ? "no error recorded" : g_err;
  return t_err.c_str();
}

std::vector<usap_phase>& engine_phase_log(usap_engine* e) { return e->phase_log; }
const usap_graph* engine_host_graph(usap_engine* e) {
  if (refresh_host_graph(e) != USAP_OK) return nullptr;
  return &e->hg;
}
const usap_params* engine_params(const usap_engine* e) { return &e->p; }
int32_t engine_rebuilds(const usap_engine* e) { return e->rebuilds; }
// Growth counters, read on the engine's own stream (never the legacy null
// stream, which other engines' threads must not see).
static bool read_grow_state(usap_engine* e, GrowState* g) {
  if (hipMemcpyAsync(g, e->M.gst, sizeof(GrowState), hipMemcpyDeviceToHost, e->stream) !=
      hipSuccess)
    return false;
  return hipStreamSynchronize(e->stream) == hipSuccess;
}
void engine_grow_debug(usap_engine* e, long long out[4]) {
  out[0] = out[1] = out[2] = out[3] = 0;
  GrowState g{};
  if (!read_grow_state(e, &g)) return;
  out[0] = (long long)g.grows;
  out[1] = (long long)g.compactions;
  out[2] = (long long)g.full;
  out[3] = (long long)g.nospace_events;
}
int64_t engine_grows(usap_engine* e) {
  GrowState g{};
  if (!read_grow_state(e, &g)) return 0;
  return (int64_t)g.grows;
}
void engine_set_error(usap_engine* e, const char* msg) { e->err = msg; }
}  // namespace usap

extern "C" {

void usap_default_params(usap_params* p) {
  p->K = 10;
  p->reduction_rate = 0.5;
  p->max_nodal_itr = 100;
  p->num_batch_nodal_update = 4;
  p->nodal_batch = 0;
  p->thr1 = 5e-4;
  p->thr2 = 1e-4;
  p->avg_window = 3;
  p->dfs_depth = 20;
  p->init = USAP_INIT_SCC;
  p->ds_mode = USAP_DS_REF;
  p->seed = 0;
  p->device = 0;
  p->verbose = 0;
  p->slab_slack = 2;
  p->graphs = 0;
}

const char* usap_last_error(const usap_engine* e) {
  if (e) return e->err.c_str();
  return usap::detached_error();
}

int usap_destroy(usap_engine* e) {
  if (!e) return USAP_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  dfree(e, e->out_ptr);
  dfree(e, e->in_ptr);
  dfree(e, e->out_idx);
  dfree(e, e->in_idx);
  dfree(e, e->out_w);
  dfree(e, e->in_w);
  dfree(e, e->out_src);
  for (int i = 0; i < 2; ++i) {
    dfree(e, e->pair_key[i]);
    dfree(e, e->pair_val[i]);
  }
  dfree(e, e->pair_tmp);
  dfree(e, e->lam);
  dfree(e, e->labels);
  dfree(e, e->d_out);
  dfree(e, e->d_in);
  dfree(e, e->ecnt_out);
  dfree(e, e->ecnt_in);
  dfree(e, e->bsize);
  dfree(e, e->bl_off);
  dfree(e, e->bl_cur);
  dfree(e, e->bl_nodes);
  for (DevSpM* m : {&e->M}) {
    dfree(e, m->off);
    dfree(e, m->cap);
    dfree(e, m->nkeys);
    dfree(e, m->kv);
    dfree(e, m->bs);
    dfree(e, m->gpend);
    dfree(e, m->ovf);
    dfree(e, m->gst);
  }
  dfree(e, e->tmpM);
  dfree(e, e->scan_tmp);
  dfree(e, e->best_s);
  dfree(e, e->best_hi);
  dfree(e, e->best_lo);
  dfree(e, e->map);
  for (int i = 0; i < 2; ++i) {
    dfree(e, e->sort_key[i]);
    dfree(e, e->sort_idx[i]);
  }
  dfree(e, e->sort_tmp);
  dfree(e, e->prop_s);
  dfree(e, e->prop_hi);
  dfree(e, e->prop_lo);
  dfree(e, e->prop_inf);
  dfree(e, e->prop_acc);
  dfree(e, e->part);
  dfree(e, e->status);
  dfree(e, e->sweep);
  dfree(e, e->prop_r);
  dfree(e, e->prop_u);
  dfree(e, e->aq_mark);
  dfree(e, e->aq_acc);
  dfree(e, e->build_flags);
  dfree(e, e->uni);
  if (e->gexec) {
    std::lock_guard<std::mutex> lk(g_hip_mu);
    (void)hipGraphExecDestroy(e->gexec);
  }
  dfree(e, e->tstamp);
  pin_put(e->h_sweep, sizeof(Sweep));
  for (int i = 0; i < 2; ++i) pin_put(e->h_chunk[i], sizeof(Sweep));
  dfree(e, e->big_off);
  dfree(e, e->big_tab);
  dfree(e, e->ent_part);
  pin_put(e->h_ent_part, sizeof(long) * 2 * kEntropyBlocks);
  for (int i = 0; i < kNumSnap; ++i) dfree(e, e->snap_labels[i]);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  {
    std::lock_guard<std::mutex> lk(g_hip_mu);
    for (hipEvent_t& ev : e->ev)
      if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t& ev : e->chunk_ev)
      if (ev) (void)hipEventDestroy(ev);
    if (e->stream && e->own_stream) (void)hipStreamDestroy(e->stream);
  }
  delete e;
  return USAP_OK;
}

}  // extern "C"

namespace {

// Checks a graph (CSR/CSC consistent, ids in range, weights >= 0, total
// weight < 2^31 since M counts and degrees are int32); *wtot = total weight.
int check_graph(const usap_graph* g, int64_t* wtot) {
  if (!g) return USAP_EINVAL;
  if (g->N < 1 || g->E < 0 || !g->out_ptr || !g->in_ptr || (g->E > 0 && (!g->out_idx || !g->in_idx)))
    return USAP_EINVAL;
  if ((g->out_w == nullptr) != (g->in_w == nullptr)) return USAP_EINVAL;
  const int32_t N = g->N;
  const int64_t E = g->E;
  if (g->out_ptr[0] != 0 || g->out_ptr[N] != E || g->in_ptr[0] != 0 || g->in_ptr[N] != E)
    return USAP_EINVAL;
  int64_t w = E;
  if (g->out_w) {
    w = 0;
    for (int64_t i = 0; i < E; ++i) {
      if (g->out_w[i] < 0) return USAP_EINVAL;
      w += g->out_w[i];
    }
  }
  for (int64_t i = 0; i < E; ++i)
    if (g->out_idx[i] < 0 || g->out_idx[i] >= N || g->in_idx[i] < 0 || g->in_idx[i] >= N)
      return USAP_EINVAL;
  if (w >= (int64_t)1 << 31) return USAP_EINVAL;
  *wtot = w;
  return USAP_OK;
}

// lambda[0..maxv] from one table shared by the process's engines (computed
// again only when a graph needs more entries: batches of graphs would
// otherwise spend most of their host time in log()).
void lambda_range(int64_t lo, int64_t hi, int64_t* out) {
  static std::mutex mu;
  static std::vector<int64_t> tab;
  std::lock_guard<std::mutex> lk(mu);
  if ((int64_t)tab.size() <= hi) {
    tab.assign((size_t)hi + 1, 0);
    usap_lambda_table(hi, tab.data());
  }
  std::memcpy(out, tab.data() + lo, sizeof(int64_t) * (size_t)(hi - lo + 1));
}

void lambda_prefix(int64_t maxv, int64_t* out) { lambda_range(0, maxv, out); }

#define TRY(x)          \
  do {                  \
    int _rc = (x);      \
    if (_rc) return _rc; \
  } while (0)

// The graph's edges on the device: host copy, CSR/CSC, edge sources, the
// lambda table over [0, wtot] and the global tables of large nodes.  Frees
// the previous graph's arrays first (same N).
int load_graph(usap_engine* e, const usap_graph* g, int64_t wtot) {
  const int32_t N = g->N;
  const int64_t E = g->E;
  e->N = N;
  e->E = E;
  e->weighted = g->out_w != nullptr;
  e->h_out_ptr.assign(g->out_ptr, g->out_ptr + N + 1);
  e->h_in_ptr.assign(g->in_ptr, g->in_ptr + N + 1);
  e->h_out_idx.assign(g->out_idx, g->out_idx + E);
  e->h_in_idx.assign(g->in_idx, g->in_idx + E);
  e->h_out_w.clear();
  e->h_in_w.clear();
  if (e->weighted) {
    e->h_out_w.assign(g->out_w, g->out_w + E);
    e->h_in_w.assign(g->in_w, g->in_w + E);
  }
  e->Wtot = wtot;
  e->hg.N = N;
  e->hg.E = E;
  e->hg.out_ptr = e->h_out_ptr.data();
  e->hg.in_ptr = e->h_in_ptr.data();
  e->hg.out_idx = e->h_out_idx.data();
  e->hg.in_idx = e->h_in_idx.data();
  e->hg.out_w = e->weighted ? e->h_out_w.data() : nullptr;
  e->hg.in_w = e->weighted ? e->h_in_w.data() : nullptr;
  for (int32_t** q : {&e->out_idx, &e->in_idx, &e->out_src, &e->out_w, &e->in_w, &e->big_tab})
    dfree(e, *q);
  for (int64_t** q : {&e->out_ptr, &e->in_ptr, &e->lam, &e->big_off}) dfree(e, *q);
  TRY(dalloc(e, &e->out_ptr, N + 1));
  TRY(dalloc(e, &e->in_ptr, N + 1));
  TRY(dalloc(e, &e->out_idx, E));
  TRY(dalloc(e, &e->in_idx, E));
  TRY(dalloc(e, &e->out_src, E));
  if (e->weighted) {
    TRY(dalloc(e, &e->out_w, E));
    TRY(dalloc(e, &e->in_w, E));
  }
  TRY(dalloc(e, &e->lam, wtot + 1));
  e->lam_cap = wtot + 1;
  e->host_stale = false;
  std::vector<int64_t> lam((size_t)wtot + 1);
  lambda_prefix(wtot, lam.data());
  std::vector<int32_t> src((size_t)E);
  for (int32_t x = 0; x < N; ++x)
    for (int64_t j = g->out_ptr[x]; j < g->out_ptr[x + 1]; ++j) src[j] = x;
  // global-memory neighbour tables of the nodes too large for LDS
  std::vector<int64_t> big((size_t)N, -1);
  int64_t big_total = 0;
  for (int32_t x = 0; x < N; ++x) {
    const int64_t k = (g->out_ptr[x + 1] - g->out_ptr[x]) + (g->in_ptr[x + 1] - g->in_ptr[x]);
    if (k > kTabMaxDeg) {
      int64_t cap = 1;
      while (cap < 2 * k) cap <<= 1;
      big[x] = big_total;
      big_total += 3 * cap;
    }
  }
  TRY(dalloc(e, &e->big_off, N));
  TRY(dalloc(e, &e->big_tab, big_total));
  hipStream_t st = e->stream;
  HIPCHK(hipMemcpyAsync(e->out_ptr, g->out_ptr, sizeof(int64_t) * (N + 1), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(e->in_ptr, g->in_ptr, sizeof(int64_t) * (N + 1), hipMemcpyHostToDevice, st));
  if (E > 0) {
    HIPCHK(hipMemcpyAsync(e->out_idx, g->out_idx, sizeof(int32_t) * E, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(e->in_idx, g->in_idx, sizeof(int32_t) * E, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(e->out_src, src.data(), sizeof(int32_t) * E, hipMemcpyHostToDevice, st));
    if (e->weighted) {
      HIPCHK(hipMemcpyAsync(e->out_w, g->out_w, sizeof(int32_t) * E, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(e->in_w, g->in_w, sizeof(int32_t) * E, hipMemcpyHostToDevice, st));
    }
  }
  HIPCHK(hipMemcpyAsync(e->lam, lam.data(), sizeof(int64_t) * (wtot + 1), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(e->big_off, big.data(), sizeof(int64_t) * N, hipMemcpyHostToDevice, st));
  // the host vectors above are released when this returns: wait for the copies
  HIPCHK(hipStreamSynchronize(st));
  return USAP_OK;
}

// Global-memory neighbour tables of the nodes whose lists exceed the LDS
// table (kTabMaxDeg), from the host copy of the offsets.
int build_big_tables(usap_engine* e) {
  const int32_t N = e->N;
  std::vector<int64_t> big((size_t)N, -1);
  int64_t big_total = 0;
  for (int32_t x = 0; x < N; ++x) {
    const int64_t k = (e->h_out_ptr[x + 1] - e->h_out_ptr[x]) + (e->h_in_ptr[x + 1] - e->h_in_ptr[x]);
    if (k > kTabMaxDeg) {
      int64_t cap = 1;
      while (cap < 2 * k) cap <<= 1;
      big[x] = big_total;
      big_total += 3 * cap;
    }
  }
  dfree(e, e->big_off);
  dfree(e, e->big_tab);
  TRY(dalloc(e, &e->big_off, N));
  TRY(dalloc(e, &e->big_tab, big_total));
  HIPCHK(hipMemcpyAsync(e->big_off, big.data(), sizeof(int64_t) * N, hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));  // `big` is released on return
  return USAP_OK;
}

// One side (CSR or CSC) of an append: the lists at ptr/idx/w grow by the n
// edges whose node is key[i] and other end other[i] (device arrays).
// owner (nullable): also rebuild the per-slot node array (the CSR's edge
// sources).
int append_side(usap_engine* e, int64_t n, const int32_t* key, const int32_t* other,
                const int32_t* w, int64_t*& ptr, int32_t*& idx, int32_t*& wv, int32_t** owner) {
  const int32_t N = e->N;
  const int64_t E1 = e->E + n;
  hipStream_t st = e->stream;
  int32_t *cnt = nullptr, *kin = nullptr, *kout = nullptr, *vin = nullptr, *vout = nullptr;
  int64_t *deg = nullptr, *nptr = nullptr;
  int32_t *nidx = nullptr, *nw = nullptr, *nown = nullptr;
  void* tmp = nullptr;
  int rc = USAP_OK;
  auto release = [&]() {
    dfree(e, cnt);
    dfree(e, kin);
    dfree(e, kout);
    dfree(e, vin);
    dfree(e, vout);
    dfree(e, deg);
    dfree(e, tmp);
  };
  if ((rc = dalloc(e, &cnt, N)) || (rc = dalloc(e, &deg, (size_t)N + 1)) ||
      (rc = dalloc(e, &nptr, (size_t)N + 1)) || (rc = dalloc(e, &kin, n)) ||
      (rc = dalloc(e, &kout, n)) || (rc = dalloc(e, &vin, n)) || (rc = dalloc(e, &vout, n)) ||
      (rc = dalloc(e, &nidx, E1)) || (e->weighted && (rc = dalloc(e, &nw, E1))) ||
      (owner && (rc = dalloc(e, &nown, E1)))) {
    release();
    dfree(e, nptr);
    dfree(e, nidx);
    dfree(e, nw);
    dfree(e, nown);
    return rc;
  }
  int bits = 1;
  while (((int64_t)1 << bits) < (int64_t)N) ++bits;
  size_t b1 = 0, b2 = 0;
  hipError_t he = rocprim::exclusive_scan(nullptr, b1, deg, nptr, (int64_t)0, (size_t)N + 1,
                                          rocprim::plus<int64_t>(), st);
  if (he == hipSuccess)
    he = rocprim::radix_sort_pairs(nullptr, b2, kin, kout, vin, vout, (size_t)n, 0, bits, st);
  if (he == hipSuccess && !(rc = dalloc_raw(e, &tmp, std::max(b1, b2)))) {
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(blocks_for(n, 256), 16384));
    if ((he = hipMemsetAsync(cnt, 0, sizeof(int32_t) * N, st)) == hipSuccess) {
      k_count_keys<<<g, 256, 0, st>>>(n, key, cnt);
      k_new_deg<<<blocks_for((int64_t)N + 1, 256), 256, 0, st>>>(N, ptr, cnt, deg);
      size_t b = std::max(b1, b2);
      he = rocprim::exclusive_scan(tmp, b, deg, nptr, (int64_t)0, (size_t)N + 1,
                                   rocprim::plus<int64_t>(), st);
    }
    if (he == hipSuccess) he = hipMemcpyAsync(kin, key, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st);
    if (he == hipSuccess) {
      k_iota<<<g, 256, 0, st>>>(n, vin);
      size_t b = std::max(b1, b2);
      he = rocprim::radix_sort_pairs(tmp, b, kin, kout, vin, vout, (size_t)n, 0, bits, st);
    }
    if (he == hipSuccess) {
      k_move_lists<<<blocks_for(N, 256 / kWave), 256, 0, st>>>(N, ptr, nptr, idx, wv, nidx, nw, nown);
      k_place_new<<<g, 256, 0, st>>>(n, kout, vout, ptr, other, w, nidx, nw, nown);
      he = hipGetLastError();
    }
  }
  release();
  if (rc == USAP_OK && he != hipSuccess)
    rc = set_err(e, USAP_EHIP, std::string("edge append: ") + hipGetErrorString(he));
  if (rc != USAP_OK) {
    dfree(e, nptr);
    dfree(e, nidx);
    dfree(e, nw);
    dfree(e, nown);
    return rc;
  }
  dfree(e, ptr);
  dfree(e, idx);
  dfree(e, wv);
  ptr = nptr;
  idx = nidx;
  wv = nw;
  if (owner) {
    dfree(e, *owner);
    *owner = nown;
  }
  return USAP_OK;
}

// The host copy of the graph from the device lists (after appends).
int refresh_host_graph(usap_engine* e) {
  if (!e->host_stale) return USAP_OK;
  const int64_t E = e->E;
  hipStream_t st = e->stream;
  e->h_out_idx.resize((size_t)E);
  e->h_in_idx.resize((size_t)E);
  e->h_out_w.resize(e->weighted ? (size_t)E : 0);
  e->h_in_w.resize(e->weighted ? (size_t)E : 0);
  if (E > 0) {
    HIPCHK(hipMemcpyAsync(e->h_out_idx.data(), e->out_idx, sizeof(int32_t) * E, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(e->h_in_idx.data(), e->in_idx, sizeof(int32_t) * E, hipMemcpyDeviceToHost, st));
    if (e->weighted) {
      HIPCHK(hipMemcpyAsync(e->h_out_w.data(), e->out_w, sizeof(int32_t) * E, hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(e->h_in_w.data(), e->in_w, sizeof(int32_t) * E, hipMemcpyDeviceToHost, st));
    }
  }
  HIPCHK(hipStreamSynchronize(st));
  e->hg.E = E;
  e->hg.out_ptr = e->h_out_ptr.data();
  e->hg.in_ptr = e->h_in_ptr.data();
  e->hg.out_idx = e->h_out_idx.data();
  e->hg.in_idx = e->h_in_idx.data();
  e->hg.out_w = e->weighted ? e->h_out_w.data() : nullptr;
  e->hg.in_w = e->weighted ? e->h_in_w.data() : nullptr;
  e->host_stale = false;
  return USAP_OK;
}

// Appends n edges (host arrays, already checked) on the device.
int append_edges(usap_engine* e, int64_t n, const int32_t* src, const int32_t* dst,
                 const int32_t* w, int64_t wsum) {
  hipStream_t st = e->stream;
  const int64_t W1 = e->Wtot + wsum;
  int32_t *ds = nullptr, *dd = nullptr, *dw = nullptr;
  auto release = [&]() {
    dfree(e, ds);
    dfree(e, dd);
    dfree(e, dw);
  };
  TRY(dalloc(e, &ds, n));
  TRY(dalloc(e, &dd, n));
  if (w) TRY(dalloc(e, &dw, n));
  HIPCHK(hipMemcpyAsync(ds, src, sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dd, dst, sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
  if (w) HIPCHK(hipMemcpyAsync(dw, w, sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
  int rc = append_side(e, n, ds, dd, dw, e->out_ptr, e->out_idx, e->out_w, &e->out_src);
  if (rc == USAP_OK) rc = append_side(e, n, dd, ds, dw, e->in_ptr, e->in_idx, e->in_w, nullptr);
  // the host arrays are the caller's: the copies above must finish first
  const hipError_t he = hipStreamSynchronize(st);
  release();
  if (rc != USAP_OK) return rc;
  if (he != hipSuccess) return set_err(e, USAP_EHIP, std::string("edge append: ") + hipGetErrorString(he));
  e->E += n;
  e->hg.E = e->E;
  // lambda over [0, W1]: the new range only, the table grown geometrically
  if (W1 + 1 > e->lam_cap) {
    const int64_t cap = std::min<int64_t>((int64_t)1 << 31, std::max<int64_t>(W1 + 1, 2 * e->lam_cap));
    int64_t* nl = nullptr;
    TRY(dalloc(e, &nl, cap));
    HIPCHK(hipMemcpyAsync(nl, e->lam, sizeof(int64_t) * (e->Wtot + 1), hipMemcpyDeviceToDevice, st));
    std::vector<int64_t> add((size_t)(cap - 1 - e->Wtot));
    lambda_range(e->Wtot + 1, cap - 1, add.data());
    HIPCHK(hipMemcpyAsync(nl + e->Wtot + 1, add.data(), sizeof(int64_t) * add.size(),
                          hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    dfree(e, e->lam);
    e->lam = nl;
    e->lam_cap = cap;
  }
  e->Wtot = W1;
  // host offsets (big tables, driver) now; lists on demand
  e->h_out_ptr.resize((size_t)e->N + 1);
  e->h_in_ptr.resize((size_t)e->N + 1);
  HIPCHK(hipMemcpyAsync(e->h_out_ptr.data(), e->out_ptr, sizeof(int64_t) * (e->N + 1), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(e->h_in_ptr.data(), e->in_ptr, sizeof(int64_t) * (e->N + 1), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  e->hg.out_ptr = e->h_out_ptr.data();
  e->hg.in_ptr = e->h_in_ptr.data();
  e->host_stale = true;
  TRY(build_big_tables(e));
  const int64_t ocap = std::min<int64_t>(kOvfCap, 2 * e->E + 4096);
  if (ocap > e->M.ovf_cap) {  // the overflow buffer follows the edge count
    dfree(e, e->M.ovf);
    e->M.ovf_cap = (int32_t)ocap;
    TRY(dalloc(e, &e->M.ovf, (size_t)4 * e->M.ovf_cap));
  }
  return USAP_OK;
}

// Device arrays sized by N only (labels, degrees, row tables, batch buffers).
int alloc_state(usap_engine* e) {
  const int32_t N = e->N;
  TRY(dalloc(e, &e->labels, N));
  TRY(dalloc(e, &e->d_out, N));
  TRY(dalloc(e, &e->d_in, N));
  TRY(dalloc(e, &e->ecnt_out, N));
  TRY(dalloc(e, &e->ecnt_in, N));
  TRY(dalloc(e, &e->bsize, N));
  TRY(dalloc(e, &e->bl_off, N));
  TRY(dalloc(e, &e->bl_cur, N));
  TRY(dalloc(e, &e->bl_nodes, N));
  for (DevSpM* m : {&e->M}) {
    TRY(dalloc(e, &m->off, (size_t)N + 1));
    TRY(dalloc(e, &m->cap, N));
    TRY(dalloc(e, &m->nkeys, N));
    TRY(dalloc(e, &m->gpend, N));
    m->ovf_cap = (int32_t)std::min<int64_t>(kOvfCap, 2 * e->E + 4096);
    TRY(dalloc(e, &m->ovf, (size_t)4 * m->ovf_cap));
    TRY(dalloc(e, &m->gst, 1));
  }
  TRY(dalloc(e, &e->tmpM, (size_t)N + 1));
  TRY(dalloc(e, &e->best_s, N));
  TRY(dalloc(e, &e->best_hi, N));
  TRY(dalloc(e, &e->best_lo, N));
  TRY(dalloc(e, &e->map, N));
  TRY(dalloc(e, &e->prop_s, N));
  TRY(dalloc(e, &e->prop_hi, N));
  TRY(dalloc(e, &e->prop_lo, N));
  TRY(dalloc(e, &e->prop_inf, N));
  TRY(dalloc(e, &e->prop_acc, N));
  TRY(dalloc(e, &e->status, 1));
  TRY(dalloc(e, &e->part, blocks_for(N, kNodalWaves)));
  TRY(dalloc(e, &e->sweep, 1));
  TRY(dalloc(e, &e->prop_r, N));
  TRY(dalloc(e, &e->build_flags, 3));
  TRY(dalloc(e, &e->ent_part, 2 * kEntropyBlocks));
  e->h_sweep = (Sweep*)pin_get(sizeof(Sweep));
  e->h_ent_part = (long*)pin_get(sizeof(long) * 2 * kEntropyBlocks);
  for (int i = 0; i < 2; ++i) e->h_chunk[i] = (Sweep*)pin_get(sizeof(Sweep));
  if (!e->h_sweep || !e->h_ent_part || !e->h_chunk[0] || !e->h_chunk[1])
    return set_err(e, USAP_ENOMEM, "hipHostMalloc");
  hipStream_t st = e->stream;
  HIPCHK(hipMemsetAsync(e->status, 0, sizeof(Status), st));
  HIPCHK(hipMemsetAsync(e->sweep, 0, sizeof(Sweep), st));
  HIPCHK(hipMemsetAsync(e->M.gst, 0, sizeof(GrowState), st));
  std::memset(e->h_sweep, 0, sizeof(Sweep));
  {
    std::lock_guard<std::mutex> lk(g_hip_mu);
    for (hipEvent_t& ev : e->ev) HIPCHK(hipEventCreate(&ev));
    for (hipEvent_t& ev : e->chunk_ev) HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  return USAP_OK;
}

#undef TRY

}  // namespace

extern "C" {

int usap_create(const usap_graph* g, const usap_params* p, usap_engine** out) {
  return usap::create_in(nullptr, g, p, out);
}

}  // extern "C"

namespace usap {

// A host worker's device context (see engine_internal.hpp).
struct WorkerCtx {
  int32_t device = 0;
  hipStream_t stream = nullptr;
  hipMemPool_t pool = nullptr;
};

// Worker contexts are kept for the life of the process: a released one goes
// back to a free list (its stream synchronised, its pool trimmed to what is
// in use) and the next worker on that device takes it, so repeated batches
// (the autotuner's trials, one batch after another) create no further
// streams or pools and never destroy one while other workers run.
std::mutex g_ctx_mu;
std::vector<WorkerCtx*> g_ctx_free;

int worker_ctx_create(int32_t device, WorkerCtx** out) {
  if (!out) return USAP_EINVAL;
  *out = nullptr;
  if (hipSetDevice(device) != hipSuccess) return USAP_EHIP;
  {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    for (size_t i = 0; i < g_ctx_free.size(); ++i)
      if (g_ctx_free[i]->device == device) {
        *out = g_ctx_free[i];
        g_ctx_free.erase(g_ctx_free.begin() + (long)i);
        return USAP_OK;
      }
  }
  WorkerCtx* w = new (std::nothrow) WorkerCtx();
  if (!w) return USAP_ENOMEM;
  w->device = device;
  std::lock_guard<std::mutex> lk(g_hip_mu);
  if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess) {
    delete w;
    return USAP_EHIP;
  }
  hipMemPoolProps props{};
  props.allocType = hipMemAllocationTypePinned;
  props.location.type = hipMemLocationTypeDevice;
  props.location.id = device;
  if (hipMemPoolCreate(&w->pool, &props) != hipSuccess) {
    (void)hipStreamDestroy(w->stream);
    delete w;
    return USAP_EHIP;
  }
  uint64_t keep = UINT64_MAX;  // freed blocks stay in the pool for the next graph
  (void)hipMemPoolSetAttribute(w->pool, hipMemPoolAttrReleaseThreshold, &keep);
  *out = w;
  return USAP_OK;
}

void worker_ctx_destroy(WorkerCtx* w) {
  if (!w) return;
  (void)hipSetDevice(w->device);
  (void)hipStreamSynchronize(w->stream);
  {
    std::lock_guard<std::mutex> lk(g_hip_mu);
    (void)hipMemPoolTrimTo(w->pool, 0);  // the blocks its engines freed go back to the device
  }
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  g_ctx_free.push_back(w);
}

int create_in(WorkerCtx* ctx, const usap_graph* g, const usap_params* p, usap_engine** out) {
  if (!g || !out) return USAP_EINVAL;
  *out = nullptr;
  int64_t wtot = 0;
  if (int rc = check_graph(g, &wtot)) return rc;
  usap_engine* e = new (std::nothrow) usap_engine();
  if (!e) return USAP_ENOMEM;
  if (p)
    e->p = *p;
  else
    usap_default_params(&e->p);
  if (e->p.K < 1 || e->p.reduction_rate <= 0 || e->p.reduction_rate >= 1 || e->p.avg_window < 2 ||
      e->p.max_nodal_itr < 1 || e->p.slab_slack < 1 || e->p.slab_slack > 16) {
    delete e;
    return USAP_EINVAL;
  }
  e->N = g->N;
  e->E = g->E;
  hipError_t hst = hipSetDevice(e->p.device);
  if (hst != hipSuccess) {
    const int rc = set_err(e, USAP_EHIP, std::string("hipSetDevice: ") + hipGetErrorString(hst));
    usap_destroy(e);
    return rc;
  }
  if (ctx) {
    if (ctx->device != e->p.device) {
      delete e;
      return USAP_EINVAL;
    }
    e->stream = ctx->stream;
    e->pool = ctx->pool;
    e->own_stream = false;
  } else {
    prepare_pool(e->p.device);
    {
      std::lock_guard<std::mutex> lk(g_hip_mu);
      hst = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    }
    if (hst != hipSuccess) {
      e->stream = nullptr;
      usap_destroy(e);
      return USAP_EHIP;
    }
  }
  int rc = alloc_state(e);
  if (rc == USAP_OK) rc = load_graph(e, g, wtot);
  if (rc != USAP_OK) {
    usap_destroy(e);
    return rc;
  }
  *out = e;
  return USAP_OK;
}

}  // namespace usap

extern "C" {

int usap_set_graph(usap_engine* e, const usap_graph* g) {
  if (!e) return USAP_EINVAL;
  int64_t wtot = 0;
  if (int rc = check_graph(g, &wtot)) return rc;
  if (g->N != e->N) return set_err(e, USAP_EINVAL, "usap_set_graph: node count differs");
  HIPCHK(hipSetDevice(e->p.device));
  HIPCHK(hipStreamSynchronize(e->stream));
  const int64_t cap = std::min<int64_t>(kOvfCap, 2 * g->E + 4096);
  if (cap > e->M.ovf_cap) {  // the overflow buffer follows the edge count
    dfree(e, e->M.ovf);
    e->M.ovf_cap = (int32_t)cap;
    if (int rc = dalloc(e, &e->M.ovf, (size_t)4 * e->M.ovf_cap)) return rc;
  }
  if (int rc = load_graph(e, g, wtot)) return rc;
  e->merge_B = -1;
  return e->has_labels ? rebuild_M(e) : USAP_OK;
}

int usap_append_edges(usap_engine* e, int64_t n, const int32_t* src, const int32_t* dst,
                      const int32_t* w) {
  if (!e || n < 0 || (n > 0 && (!src || !dst))) return USAP_EINVAL;
  if (w != nullptr && !e->weighted)
    return set_err(e, USAP_EINVAL, "usap_append_edges: weights for an unweighted graph");
  if (n == 0) return USAP_OK;
  int64_t wsum = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (src[i] < 0 || src[i] >= e->N || dst[i] < 0 || dst[i] >= e->N)
      return set_err(e, USAP_EINVAL, "usap_append_edges: node id out of range");
    const int64_t wi = w ? w[i] : 1;
    if (wi < 0) return set_err(e, USAP_EINVAL, "usap_append_edges: negative weight");
    wsum += wi;
  }
  if (e->Wtot + wsum >= (int64_t)1 << 31)
    return set_err(e, USAP_EINVAL, "usap_append_edges: total weight reaches 2^31");
  HIPCHK(hipSetDevice(e->p.device));
  if (int rc = append_edges(e, n, src, dst, w, wsum)) return rc;
  e->merge_B = -1;
  return e->has_labels ? rebuild_M(e) : USAP_OK;
}

int usap_create_from_edges(int32_t N, int64_t E, const int32_t* src, const int32_t* dst,
                           const int32_t* w, const usap_params* p, usap_engine** out) {
  if (!out || N < 1 || E < 0 || (E > 0 && (!src || !dst))) return USAP_EINVAL;
  *out = nullptr;
  std::vector<int64_t> zero((size_t)N + 1, 0);
  const int32_t dummy_w = 1;
  usap_graph g{};
  g.N = N;
  g.E = 0;
  g.out_ptr = zero.data();
  g.in_ptr = zero.data();
  g.out_w = w ? &dummy_w : nullptr;  // weighted iff weights are given
  g.in_w = g.out_w;
  usap_engine* e = nullptr;
  if (int rc = usap_create(&g, p, &e)) return rc;
  if (int rc = usap_append_edges(e, E, src, dst, w)) {
    usap::set_detached_error(std::string("usap_create_from_edges: ") + e->err);
    usap_destroy(e);
    return rc;
  }
  *out = e;
  return USAP_OK;
}

int usap_set_labels(usap_engine* e, const int32_t* labels, int32_t B) {
  if (!e || !labels || B < 1 || B > e->N) return USAP_EINVAL;
  for (int32_t x = 0; x < e->N; ++x)
    if (labels[x] < 0 || labels[x] >= B) return set_err(e, USAP_EINVAL, "label out of range");
  HIPCHK(hipMemcpyAsync(e->labels, labels, sizeof(int32_t) * e->N, hipMemcpyHostToDevice, e->stream));
  e->B = B;
  e->has_labels = true;
  e->merge_B = -1;
  return rebuild_M(e);
}

int usap_get_labels(usap_engine* e, int32_t* out) {
  if (int rc = check_engine(e)) return rc;
  if (!out) return USAP_EINVAL;
  HIPCHK(hipMemcpyAsync(out, e->labels, sizeof(int32_t) * e->N, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return USAP_OK;
}

int usap_num_blocks(usap_engine* e, int32_t* B) {
  if (int rc = check_engine(e)) return rc;
  if (!B) return USAP_EINVAL;
  *B = e->B;
  return USAP_OK;
}

int usap_get_degrees(usap_engine* e, int32_t* d_out, int32_t* d_in) {
  if (int rc = check_engine(e)) return rc;
  if (d_out)
    HIPCHK(hipMemcpyAsync(d_out, e->d_out, sizeof(int32_t) * e->B, hipMemcpyDeviceToHost, e->stream));
  if (d_in)
    HIPCHK(hipMemcpyAsync(d_in, e->d_in, sizeof(int32_t) * e->B, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return USAP_OK;
}

int usap_get_M_coo(usap_engine* e, int64_t cap, int32_t* rows, int32_t* cols, int32_t* vals,
                   int64_t* nnz) {
  if (int rc = check_engine(e)) return rc;
  if (!nnz) return USAP_EINVAL;
  const int32_t B = e->B;
  hipStream_t st = e->stream;
  GrowState gs{};
  HIPCHK(hipMemcpyAsync(&gs, e->M.gst, sizeof(GrowState), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const int64_t used = std::min<int64_t>((int64_t)gs.used, e->M.P);
  std::vector<int64_t> off((size_t)B + 1);
  std::vector<int32_t> capv((size_t)B);
  std::vector<Slot> kv((size_t)used);
  HIPCHK(hipMemcpyAsync(off.data(), e->M.off, sizeof(int64_t) * (B + 1), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(capv.data(), e->M.cap, sizeof(int32_t) * B, hipMemcpyDeviceToHost, st));
  if (used > 0) {
    HIPCHK(hipMemcpyAsync(kv.data(), e->M.kv, sizeof(Slot) * used, hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  std::vector<std::pair<int32_t, int32_t>> row;
  int64_t k = 0;
  for (int32_t a = 0; a < B; ++a) {
    row.clear();
    for (int64_t j = 0; j < capv[a]; ++j) {
      const int32_t c = capv[a] == B ? (int32_t)j : kv[off[a] + j].k;
      const int32_t v = kv[off[a] + j].o;  // M[a][c]
      if (c >= 0 && v != 0) row.emplace_back(c, v);
    }
    std::sort(row.begin(), row.end());
    for (auto& cv : row) {
      if (k < cap && rows && cols && vals) {
        rows[k] = a;
        cols[k] = cv.first;
        vals[k] = cv.second;
      }
      ++k;
    }
  }
  *nnz = k;
  return USAP_OK;
}

}  // extern "C"

namespace {
// k_entropy over the pool in use: about four slots per thread, at most
// kEntropyBlocks workgroups (a scan of 1 GB pools at B = 250k is bound by the
// loads in flight).  Returns the grid (the number of partials).
int launch_entropy(usap_engine* e, long* part = nullptr) {
  const int64_t work = std::max<int64_t>(e->M.P, e->B);
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(kEntropyBlocks, blocks_for(work, 1024)));
  k_entropy<<<grid, 256, 0, e->stream>>>(e->M.kv, e->M.gst, e->d_out, e->d_in, e->B, e->lam,
                                         e->M.P, part ? part : e->ent_part);
  return grid;
}

// ---- allocation-free MDL -----------------------------------------------------
// usap_entropy_noalloc keeps k_entropy's per-workgroup partials in this
// file-scope array instead of the engine's buffer: kEntropyBlocks partials of
// 16 bytes (64 KiB), one copy per GPU the library runs on, held until the
// process ends.  Host code never passes the symbol itself: each device's copy
// is reached through the address hipGetSymbolAddress returns for it on that
// device.
constexpr int kScratchParts = kEntropyBlocks;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
long* g_scratch_ptr[kMaxDevices] = {};

// The MDL from k_entropy's nparts partials, into the caller's device memory
// (S: model + data term; fx: the data term's int128 words, hi then lo): the
// same integer sum and IEEE operations as usap_entropy on the host.
__global__ __launch_bounds__(256) void k_entropy_out(const long* __restrict__ part, int32_t nparts,
                                                     double model, double* S, int64_t* fx) {
  __shared__ long s_hi[4];
  __shared__ unsigned long s_lo[4];
  i128 t = 0;
  for (int32_t i = threadIdx.x; i < nparts; i += blockDim.x)
    t += (i128)(((u128)(unsigned long)part[2 * i] << 64) | (u128)(unsigned long)part[2 * i + 1]);
  t = wave_sum128(t);
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) {
    s_hi[w] = (long)(t >> 64);
    s_lo[w] = (unsigned long)(u128)t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    i128 tot = 0;
    for (int i = 0; i < (int)(blockDim.x / kWave); ++i)
      tot += (i128)(((u128)(unsigned long)s_hi[i] << 64) | (u128)s_lo[i]);
    const i128 data = -tot;
    if (S) *S = model + fx_to_double_dev(data);
    if (fx) {
      fx[0] = (int64_t)(data >> 64);
      fx[1] = (int64_t)(uint64_t)(u128)data;
    }
  }
}

// This device's copy of `scratch`: looked up once per device, with the
// device current, on the first allocation-free call there.
int scratch_on(usap_engine* e, long** out) {
  const int dev = e->p.device;
  if (dev < 0 || dev >= kMaxDevices)
    return set_err(e, USAP_EINVAL, "usap_entropy_noalloc: device ordinal beyond 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = -1;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (st != hipSuccess || !p)
      return set_err(e, USAP_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(st));
    g_scratch_ptr[dev] = static_cast<long*>(p);
  }
  *out = g_scratch_ptr[dev];
  return USAP_OK;
}
}  // namespace

extern "C" {

int usap_entropy(usap_engine* e, double* S, int64_t* hi, uint64_t* lo) {
  if (int rc = check_engine(e)) return rc;
  hipStream_t st = e->stream;
  ev_rec(e, 0);
  const int grid = launch_entropy(e);
  HIPCHK(hipGetLastError());
  ev_rec(e, 1);
  HIPCHK(hipMemcpyAsync(e->h_ent_part, e->ent_part, sizeof(long) * 2 * grid, hipMemcpyDeviceToHost,
                        st));
  HIPCHK(hipStreamSynchronize(st));
  prof_add(e, USAP_K_ENTROPY, ev_ms(e, 0, 1), 1, 8.0 * e->B + 16.0 * grid);
  i128 t = 0;
  for (int i = 0; i < grid; ++i)
    t += mk128(e->h_ent_part[2 * i], (unsigned long)e->h_ent_part[2 * i + 1]);
  const i128 data = -t;
  if (hi) *hi = (int64_t)(data >> 64);
  if (lo) *lo = (uint64_t)(u128)data;
  if (S) *S = usap_model_S(e->N, e->E, e->B) + fx_to_double(data);
  return USAP_OK;
}

int usap_entropy_noalloc(usap_engine* e, double* S_dev, int64_t* fx_dev) {
  if (int rc = check_engine(e)) return rc;
  if (!S_dev && !fx_dev) return USAP_EINVAL;
  long* part = nullptr;
  if (int rc = scratch_on(e, &part)) return rc;
  // k_entropy is grid-strided over the pool with at most kScratchParts
  // workgroups, so every engine state fits the scratch
  const int64_t need = std::min<int64_t>(kEntropyBlocks, blocks_for(std::max<int64_t>(e->M.P, e->B), 1024));
  if (need > kScratchParts)
    return set_err(e, USAP_EINVAL, "usap_entropy_noalloc: more partials than the 64 KiB scratch holds");
  const int grid = launch_entropy(e, part);
  k_entropy_out<<<1, 256, 0, e->stream>>>(part, grid, usap_model_S(e->N, e->E, e->B), S_dev, fx_dev);
  HIPCHK(hipGetLastError());
  return USAP_OK;
}

int usap_merge_propose(usap_engine* e, int32_t K, uint64_t stream, int32_t* best_s,
                       double* best_dS) {
  if (int rc = check_engine(e)) return rc;
  if (K < 1) return USAP_EINVAL;
  const int32_t B = e->B;
  MergeArgs a;
  a.M = view(e->M, B);
  a.d_out = e->d_out;
  a.d_in = e->d_in;
  a.lam = e->lam;
  a.B = B;
  a.K = K;
  a.seed = e->p.seed;
  a.stream = stream;
  a.uni = e->uni;
  if (e->uni) {  // the round takes draws uni_pos + r*K + idx
    const int64_t need = (int64_t)B * K;
    if (need > e->uni_n - e->uni_pos)
      return set_err(e, USAP_ESTATE, "pre-drawn uniform stream exhausted: a merge round needs " +
                                         std::to_string(need) + " draws, " +
                                         std::to_string(e->uni_n - e->uni_pos) + " left");
    a.stream = (uint64_t)e->uni_pos;
  }
  a.best_s = e->best_s;
  a.best_hi = e->best_hi;
  a.best_lo = e->best_lo;
  const int nw = std::min(K, kMergeWaves);
  ev_rec(e, 0);
  k_merge<<<B, nw * kWave, 0, e->stream>>>(a);
  HIPCHK(hipGetLastError());
  if (e->uni) e->uni_pos += (int64_t)B * K;
  ev_rec(e, 1);
  if (e->prof) {  // the reference algorithm's bytes: 2 (4 (B+1) + 8 nnz(M)) + 24 B
    k_nnz_bytes<<<(unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, blocks_for(e->M.P, 256))),
                  256, 0, e->stream>>>(e->M.kv, e->M.P, e->aq_acc + 2 * kAqSlots);
    HIPCHK(hipGetLastError());
    e->kp[USAP_K_MERGE].model_bytes += 8.0 * ((double)B + 1.0) + 24.0 * (double)B;
  }
  e->merge_B = B;
  if (best_s || best_dS) {
    std::vector<long> h((size_t)B);
    std::vector<unsigned long> l((size_t)B);
    if (best_s)
      HIPCHK(hipMemcpyAsync(best_s, e->best_s, sizeof(int32_t) * B, hipMemcpyDeviceToHost, e->stream));
    if (best_dS) {
      HIPCHK(hipMemcpyAsync(h.data(), e->best_hi, sizeof(long) * B, hipMemcpyDeviceToHost, e->stream));
      HIPCHK(hipMemcpyAsync(l.data(), e->best_lo, sizeof(long) * B, hipMemcpyDeviceToHost, e->stream));
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    if (best_dS)
      for (int32_t r = 0; r < B; ++r) best_dS[r] = fx_to_double(mk128(h[r], l[r]));
  } else if (e->prof) {  // the kernel time is read below; else the caller's next step waits
    HIPCHK(hipStreamSynchronize(e->stream));
  }
  prof_add(e, USAP_K_MERGE, ev_ms(e, 0, 1), 1, 0.0);
  return USAP_OK;
}

int usap_merge_result_fx(usap_engine* e, int32_t* best_s, int64_t* hi, uint64_t* lo) {
  if (int rc = check_engine(e)) return rc;
  if (e->merge_B != e->B) return set_err(e, USAP_ESTATE, "no merge proposals for this state");
  const int32_t B = e->B;
  if (best_s) HIPCHK(hipMemcpyAsync(best_s, e->best_s, sizeof(int32_t) * B, hipMemcpyDeviceToHost, e->stream));
  if (hi) HIPCHK(hipMemcpyAsync(hi, e->best_hi, sizeof(int64_t) * B, hipMemcpyDeviceToHost, e->stream));
  if (lo) HIPCHK(hipMemcpyAsync(lo, e->best_lo, sizeof(uint64_t) * B, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return USAP_OK;
}

int usap_merge_apply(usap_engine* e, int32_t B_to_merge, int32_t* B_new) {
  if (int rc = check_engine(e)) return rc;
  const int32_t B = e->B;
  if (e->merge_B != B) return set_err(e, USAP_ESTATE, "merge_apply without merge_propose");
  if (B_to_merge < 0 || B_to_merge >= B) return set_err(e, USAP_EINVAL, "B_to_merge out of range");
  std::vector<int32_t> best((size_t)B), bsize((size_t)B), order((size_t)B);
  hipStream_t st = e->stream;
  // order by (dS, block id) on the device: the reference sorts by dS with an
  // unstable sort (graph_parallel.hpp:1230)
  if (!e->sort_idx[0]) {
    for (int i = 0; i < 2; ++i) {
      if (int rc = dalloc(e, &e->sort_key[i], e->N)) return rc;
      if (int rc = dalloc(e, &e->sort_idx[i], e->N)) return rc;
    }
  }
  size_t bytes = 0;
  HIPCHK(rocprim::radix_sort_pairs(nullptr, bytes, e->sort_key[0], e->sort_key[1], e->sort_idx[0],
                                   e->sort_idx[1], (size_t)B, 0, 64, st));
  if (bytes > e->sort_tmp_bytes) {
    dfree(e, e->sort_tmp);
    if (int rc = dalloc_raw(e, &e->sort_tmp, bytes)) return rc;
    e->sort_tmp_bytes = bytes;
  }
  k_sort_init<<<blocks_for(B, 256), 256, 0, st>>>(B, e->best_lo, e->sort_key[0], e->sort_idx[0]);
  bytes = e->sort_tmp_bytes;
  HIPCHK(rocprim::radix_sort_pairs(e->sort_tmp, bytes, e->sort_key[0], e->sort_key[1],
                                   e->sort_idx[0], e->sort_idx[1], (size_t)B, 0, 64, st));
  k_sort_hi<<<blocks_for(B, 256), 256, 0, st>>>(B, e->best_hi, e->sort_idx[1], e->sort_key[0]);
  bytes = e->sort_tmp_bytes;
  HIPCHK(rocprim::radix_sort_pairs(e->sort_tmp, bytes, e->sort_key[0], e->sort_key[1],
                                   e->sort_idx[1], e->sort_idx[0], (size_t)B, 0, 64, st));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(order.data(), e->sort_idx[0], sizeof(int32_t) * B, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(best.data(), e->best_s, sizeof(int32_t) * B, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(bsize.data(), e->bsize, sizeof(int32_t) * B, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  // greedy merges: block_map of graph_parallel.hpp:1236-1257 as a union-find
  // whose root is the representative (the merge target's current block)
  std::vector<int32_t> parent((size_t)B);
  std::iota(parent.begin(), parent.end(), 0);
  auto find = [&](int32_t x) {
    int32_t r = x;
    while (parent[r] != r) r = parent[r];
    while (parent[x] != r) {
      const int32_t n = parent[x];
      parent[x] = r;
      x = n;
    }
    return r;
  };
  int32_t num = 0, counter = 0;
  while (num < B_to_merge) {
    if (counter >= B) return set_err(e, USAP_EINFEASIBLE, "carry-out ran out of candidate merges");
    const int32_t from = order[counter++];
    const int32_t to = find(best[from]);
    if (to != from) {
      parent[from] = to;
      ++num;
    }
  }
  // compaction: surviving labels ranked ascending (graph_parallel.hpp:1260-1277)
  std::vector<char> alive((size_t)B, 0);
  for (int32_t b = 0; b < B; ++b)
    if (bsize[b] > 0) alive[find(b)] = 1;
  std::vector<int32_t> rank((size_t)B, -1);
  int32_t next = 0;
  for (int32_t b = 0; b < B; ++b)
    if (alive[b]) rank[b] = next++;
  std::vector<int32_t> map((size_t)B);
  for (int32_t b = 0; b < B; ++b) map[b] = rank[find(b)] < 0 ? 0 : rank[find(b)];
  HIPCHK(hipMemcpyAsync(e->map, map.data(), sizeof(int32_t) * B, hipMemcpyHostToDevice, st));
  k_relabel<<<blocks_for(e->N, 256), 256, 0, st>>>(e->N, e->labels, e->map);
  HIPCHK(hipGetLastError());
  e->B = B - B_to_merge;
  e->merge_B = -1;
  if (B_new) *B_new = e->B;
  return rebuild_M(e);
}

}  // extern "C"

namespace {

// Profiling of a round of batch launches (slot i = the round's batch i, of
// at most `batch` nodes): every workgroup of k_nodal, k_apply and k_grow
// stamps its start and end, and each kernel's time is its first workgroup
// start to its last workgroup end, the kernel's own execution.  Events on
// these launches would add cache flushes between the kernels and slow them.
int prof_begin(usap_engine* e, size_t nslots, int32_t batch) {
  const int32_t G = (int32_t)blocks_for(batch, kNodalWaves);
  const size_t n = nslots * (4 * (size_t)G + 2 + 2 * kGrowBlocks) + 6 * nslots;
  if (e->tstamp_n < n) {
    dfree(e, e->tstamp);
    if (int rc = dalloc(e, &e->tstamp, n)) return rc;
    e->tstamp_n = n;
  }
  e->tstamp_G = G;
  k_stamp_init<<<std::min<unsigned>(blocks_for((int64_t)n, 256), 4096u), 256, 0, e->stream>>>(
      e->tstamp, (int64_t)n);
  HIPCHK(hipGetLastError());
  return USAP_OK;
}

// Adds the kernel times of the round's first `executed` batch slots to the
// profile (after the round's launches).
int prof_end(usap_engine* e, size_t executed) {
  if (executed == 0) return USAP_OK;
  const int32_t G = e->tstamp_G;
  unsigned long long* out = e->tstamp + e->tstamp_n - 6 * executed;  // past every slot's stamps
  k_stamp_reduce<<<(unsigned)(3 * executed), 256, 0, e->stream>>>(e->tstamp, G, kGrowBlocks, out);
  HIPCHK(hipGetLastError());
  std::vector<unsigned long long> t(6 * executed);
  HIPCHK(hipMemcpyAsync(t.data(), out, sizeof(unsigned long long) * 6 * executed,
                        hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  const int kid[3] = {USAP_K_NODAL, USAP_K_APPLY, USAP_K_GROW};
  for (size_t i = 0; i < executed; ++i) {
    for (int k = 0; k < 3; ++k) {
      const unsigned long long t0 = t[6 * i + 2 * k], t1 = t[6 * i + 2 * k + 1];
      if (t1 > t0) e->kp[kid[k]].ms += 1e-5 * (double)(t1 - t0);  // 100 MHz clock
      e->kp[kid[k]].launches += 1;
    }
  }
  return USAP_OK;
}

// Launches k_nodal + k_apply + k_grow for nodes [beg, end), batch bidx of the
// sweep (random stream sweep->stream0 + bidx).
int launch_batch(usap_engine* e, int32_t beg, int32_t end, int32_t bidx, size_t ev_slot,
                 bool grow = true) {
  const int32_t B = e->B;
  hipStream_t st = e->stream;
  NodalArgs a;
  a.g = gview(e);
  a.M = view(e->M, B);
  a.labels = e->labels;
  a.d_out = e->d_out;
  a.d_in = e->d_in;
  a.lam = e->lam;
  a.B = B;
  a.beg = beg;
  a.end = end;
  a.exact = e->p.ds_mode == USAP_DS_EXACT;
  a.seed = e->p.seed;
  a.bidx = bidx;
  // profiled: this slot's workgroup stamps (k_nodal: G starts + G ends;
  // k_apply: G + 1 each; k_grow: kGrowBlocks each)
  const int32_t G = e->tstamp_G;
  unsigned long long* ts =
      e->prof ? e->tstamp + ev_slot * (4 * (size_t)G + 2 + 2 * kGrowBlocks) : nullptr;
  a.tstamp = ts;
  a.tstride = G;
  a.big_off = e->big_off;
  a.big_tab = e->big_tab;
  a.prop_s = e->prop_s;
  a.prop_hi = e->prop_hi;
  a.prop_lo = e->prop_lo;
  a.prop_inf = e->prop_inf;
  a.prop_acc = e->prop_acc;
  a.prop_r = e->prop_r;
  a.prop_u = e->prof ? e->prop_u : nullptr;
  a.uni = e->uni;
  a.part = e->part;
  a.status = e->status;
  a.sweep = e->sweep;
  ApplyArgs b;
  b.g = a.g;
  b.M = a.M;
  b.labels = e->labels;
  b.d_out = e->d_out;
  b.d_in = e->d_in;
  b.bsize = e->bsize;
  b.beg = beg;
  b.end = end;
  b.prop_s = e->prop_s;
  b.prop_hi = e->prop_hi;
  b.prop_lo = e->prop_lo;
  b.prop_inf = e->prop_inf;
  b.prop_acc = e->prop_acc;
  b.prop_r = e->prop_r;
  b.big_off = e->big_off;
  b.big_tab = e->big_tab;
  const unsigned grid = blocks_for((int64_t)end - beg, kNodalWaves);
  b.part = e->part;
  b.nparts = grid;
  b.status = e->status;
  b.sweep = e->sweep;
  b.tstamp = ts ? ts + 2 * (size_t)G : nullptr;
  b.tstride = G + 1;
  GrowArgs ga;
  {
    const DevSpM& m = e->M;
    GrowSide& gs = ga.side;
    gs.off = m.off;
    gs.cap = m.cap;
    gs.nkeys = m.nkeys;
    gs.kv = m.kv;
    gs.gpend = m.gpend;
    gs.ovf = m.ovf;
    gs.ovf_cap = m.ovf_cap;
    gs.gst = m.gst;
    gs.P = m.P;
    gs.slack = m.slack;
  }
  ga.B = B;
  ga.sweep = e->sweep;
  ga.status = e->status;
  ga.tstamp = ts ? ts + 4 * (size_t)G + 2 : nullptr;
  ga.tstride = kGrowBlocks;
  if (e->prof && (int64_t)grid > G)
    return set_err(e, USAP_ESTATE, "profiled batch larger than its stamp slot");
  k_nodal<<<grid, kWave * kNodalWaves, 0, st>>>(a);
  if (e->prof) {  // the reference algorithm's bytes for this batch (M not yet changed)
    AqArgs q;
    q.g = a.g;
    q.M = a.M;
    q.beg = beg;
    q.end = end;
    q.prop_r = e->prop_r;
    q.prop_s = e->prop_s;
    q.prop_u = e->prop_u;
    q.prop_acc = e->prop_acc;
    q.mark = e->aq_mark;
    q.stamp = e->aq_stamp++;
    q.sweep = e->sweep;
    q.acc = e->aq_acc;
    k_aq<<<grid, kWave * kNodalWaves, 0, st>>>(q);
  }
  k_apply<<<grid + 1, kWave * kNodalWaves, 0, st>>>(b);  // block 0: the update step
  if (grow) k_grow<<<kGrowBlocks, 256, 0, st>>>(ga);
  HIPCHK(hipGetLastError());
  return USAP_OK;
}

// The batches [0, nb) of a sweep as one HIP graph, captured on first use
// and replayed while B, the batch size and every device pointer the batch
// kernels take (epoch) stay the same; the sweep's state (random stream,
// convergence inputs) lives in device memory, so the kernels' arguments do
// not change between sweeps.  The graph holds kernel nodes only.
int sweep_graph(usap_engine* e, int32_t batch, int32_t nb) {
  hipStream_t st = e->stream;
  if (!e->gexec || e->g_B != e->B || e->g_batch != batch || e->g_nb != nb ||
      e->g_epoch != e->epoch) {
    if (e->gexec) {
      std::lock_guard<std::mutex> lk(g_hip_mu);
      (void)hipGraphExecDestroy(e->gexec);
      e->gexec = nullptr;
    }
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = USAP_OK;
    for (int32_t i = 0; i < nb && rc == USAP_OK; ++i) {
      const int32_t beg = i * batch;
      const int32_t end = std::min<int64_t>((int64_t)beg + batch, e->N);
      rc = launch_batch(e, beg, end, i, 0);
    }
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(st, &graph);
    std::unique_lock<std::mutex> lk(g_hip_mu);
    if (rc != USAP_OK) {
      if (ce == hipSuccess && graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    HIPCHK(ce);
    const hipError_t ie = hipGraphInstantiate(&e->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    lk.unlock();
    if (ie != hipSuccess) {
      e->gexec = nullptr;
      return set_err(e, USAP_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
    e->g_B = e->B;
    e->g_batch = batch;
    e->g_nb = nb;
    e->g_epoch = e->epoch;
    e->graph_captures++;
  }
  HIPCHK(hipGraphLaunch(e->gexec, st));
  return USAP_OK;
}

// Runs the batches [first, last) of a sweep (batch nodes each) from the
// device state already in e->sweep, one host synchronisation at the end;
// rebuilds M and resumes when a batch overflowed a row slab.  Stops early
// when the sweep's convergence test fires.
constexpr int32_t kChunk = 16;  // batches per launch chunk of a sweep that may converge

int run_sweep(usap_engine* e, int32_t batch, int32_t first, int32_t last, int32_t* rebuilds) {
  hipStream_t st = e->stream;
  Sweep& hs = *e->h_sweep;
  int32_t next = first;
  for (;;) {
    if (e->p.graphs && !e->prof && next == 0) {
      if (int rc = sweep_graph(e, batch, last)) return rc;
    } else {
      if (e->prof)
        if (int rc = prof_begin(e, (size_t)(last - next), batch)) return rc;
      // A sweep whose convergence test is active stops at a batch nobody
      // knows in advance, and the batches launched after it exit at once
      // (a few us each).  Such a sweep goes out in chunks of kChunk batches,
      // one chunk ahead of the host: before the host waits for chunk c's
      // state, chunk c + 1 is queued, so the device never idles and at most
      // one chunk is launched past the stop.
      const bool chunked = !e->prof && hs.check && last - next > 2 * kChunk;
      size_t slot = 0;
      int32_t c = 0;
      for (int32_t i0 = next; i0 < last; ++c) {
        const int32_t upto = chunked ? std::min<int32_t>(last, i0 + kChunk) : last;
        for (int32_t i = i0; i < upto; ++i, ++slot) {
          const int32_t beg = i * batch;
          const int32_t end = std::min<int64_t>((int64_t)beg + batch, e->N);
          // with every row direct, the chunk's last k_grow folds the error
          // flags of all its batches (they only report inconsistent M)
          const bool grow = !e->all_direct || e->prof || i == upto - 1;
          if (int rc = launch_batch(e, beg, end, i, slot, grow)) return rc;
        }
        i0 = upto;
        if (!chunked) break;
        HIPCHK(hipMemcpyAsync(e->h_chunk[c & 1], e->sweep, sizeof(Sweep), hipMemcpyDeviceToHost, st));
        HIPCHK(hipEventRecord(e->chunk_ev[c & 1], st));
        if (c > 0) {
          HIPCHK(hipEventSynchronize(e->chunk_ev[(c - 1) & 1]));
          if (e->h_chunk[(c - 1) & 1]->stop != 0) break;
        }
      }
    }
    HIPCHK(hipMemcpyAsync(&hs, e->sweep, sizeof(Sweep), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const int32_t executed = first + hs.done - next;  // batches run in this launch round
    if (e->prof)
      if (int rc = prof_end(e, (size_t)executed)) return rc;
    if (hs.stop == 3) return set_err(e, USAP_ESTATE, "inconsistent M during a nodal batch");
    if (hs.flags & (kFlagLoad | kFlagFull)) {
      if (int rc = rebuild_M(e)) return rc;
      if (rebuilds) ++*rebuilds;
    }
    if (hs.stop != 2 || first + hs.done >= last) break;
    // resume after the rebuild
    next = first + hs.done;
    hs.stop = 0;
    hs.flags = 0;
    HIPCHK(hipMemcpyAsync(e->sweep, &hs, sizeof(Sweep), hipMemcpyHostToDevice, st));
  }
  if (e->prof) {
    e->kp[USAP_K_NODAL].bytes += (double)hs.nodal_bytes;
    e->kp[USAP_K_APPLY].bytes += (double)hs.apply_bytes;
  }
  return USAP_OK;
}

void init_sweep(usap_engine* e, uint64_t stream0, double S, double mean, double thr,
                int32_t check) {
  Sweep& hs = *e->h_sweep;
  std::memset(&hs, 0, sizeof(Sweep));
  hs.stream0 = stream0;
  hs.S = S;
  hs.mean = mean;
  hs.thr = thr;
  hs.check = check;
}

}  // namespace

extern "C" {

int usap_nodal_batch(usap_engine* e, int32_t begin, int32_t end, uint64_t stream, double* dS_sum,
                     int32_t* n_moves) {
  if (int rc = check_engine(e)) return rc;
  if (begin < 0 || end > e->N || begin > end) return USAP_EINVAL;
  if (e->uni) {  // node n takes draw uni_pos + (n - begin)
    if (end - begin > e->uni_n - e->uni_pos)
      return set_err(e, USAP_ESTATE, "pre-drawn uniform stream exhausted: a nodal batch needs " +
                                         std::to_string(end - begin) + " draws, " +
                                         std::to_string(e->uni_n - e->uni_pos) + " left");
    stream = (uint64_t)(e->uni_pos - begin);
  }
  e->last_beg = begin;
  e->last_end = end;
  init_sweep(e, stream, 0.0, 0.0, 0.0, 0);
  if (begin < end) {
    HIPCHK(hipMemcpyAsync(e->sweep, e->h_sweep, sizeof(Sweep), hipMemcpyHostToDevice, e->stream));
    // one batch of size end-begin starting at begin: index 0 of a sweep whose
    // "batch" starts at begin
    if (e->prof)
      if (int rc = prof_begin(e, 1, end - begin)) return rc;
    if (int rc = launch_batch(e, begin, end, 0, 0)) return rc;
    Sweep& hs = *e->h_sweep;
    HIPCHK(hipMemcpyAsync(&hs, e->sweep, sizeof(Sweep), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (e->prof) {
      if (int rc = prof_end(e, 1)) return rc;
      e->kp[USAP_K_NODAL].bytes += (double)hs.nodal_bytes;
      e->kp[USAP_K_APPLY].bytes += (double)hs.apply_bytes;
    }
    if (hs.stop == 3) return set_err(e, USAP_ESTATE, "inconsistent M during a nodal batch");
    if (e->uni) e->uni_pos += end - begin;
    if (hs.flags & (kFlagLoad | kFlagFull))
      if (int rc = rebuild_M(e)) return rc;
  }
  const Sweep& hs = *e->h_sweep;
  if (dS_sum) *dS_sum = hs.cum_inf ? -INFINITY : fx_to_double(mk128((long)hs.cum_hi, (unsigned long)hs.cum_lo));
  if (n_moves) *n_moves = (int32_t)hs.moves;
  return USAP_OK;
}

int usap_nodal_sweep(usap_engine* e, int32_t batch, uint64_t stream0, double S, double mean,
                     double thr, int32_t check, usap_sweep* out) {
  return usap::nodal_sweep(e, batch, stream0, &S, mean, thr, check, out);
}

int usap_set_uniforms(usap_engine* e, const double* u, size_t n) {
  if (!e) return USAP_EINVAL;
  if (n % 4 || (n && !u))
    return set_err(e, USAP_EINVAL, "usap_set_uniforms: n must be a multiple of 4 (four uniforms per draw)");
  std::vector<uint64_t> w(n);
  for (size_t i = 0; i < n; ++i) {
    if (!(u[i] >= 0.0 && u[i] < 1.0))
      return set_err(e, USAP_EINVAL, "usap_set_uniforms: u[" + std::to_string(i) + "] is not in [0, 1)");
    w[i] = (uint64_t)std::ldexp(u[i], 64);  // floor(u * 2^64): exact below 2^64
  }
  dfree(e, e->uni);
  e->uni_n = 0;
  e->uni_pos = 0;
  e->epoch++;  // a captured sweep holds the old pointer
  if (n) {
    if (int rc = dalloc(e, &e->uni, n)) return rc;
    HIPCHK(hipMemcpyAsync(e->uni, w.data(), sizeof(uint64_t) * n, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->uni_n = (int64_t)(n / 4);
  }
  return USAP_OK;
}

int usap_uniforms_used(usap_engine* e, int64_t* draws) {
  if (!e || !draws) return USAP_EINVAL;
  *draws = e->uni_pos;
  return USAP_OK;
}

}  // extern "C"

namespace usap {

// usap_nodal_sweep; with S null the sweep's starting S is the entropy of the
// current state, computed on the device between the state copy and the first
// batch (no host wait for it).
int nodal_sweep(usap_engine* e, int32_t batch, uint64_t stream0, const double* S, double mean,
                double thr, int32_t check, usap_sweep* out) {
  if (int rc = check_engine(e)) return rc;
  if (batch < 1 || !out) return USAP_EINVAL;
  double S0 = 0.0;
  if (!S && e->prof) {  // profiled: the entropy scan is timed on its own
    if (int rc = usap_entropy(e, &S0, nullptr, nullptr)) return rc;
    S = &S0;
  }
  const int32_t nb = (int32_t)(((int64_t)e->N + batch - 1) / batch);
  if (e->uni) {  // node n takes draw uni_pos + n; the sweep needs all N available
    if (e->N > e->uni_n - e->uni_pos)
      return set_err(e, USAP_ESTATE, "pre-drawn uniform stream exhausted: a nodal sweep needs " +
                                         std::to_string(e->N) + " draws, " +
                                         std::to_string(e->uni_n - e->uni_pos) + " left");
    stream0 = (uint64_t)e->uni_pos;
  }
  init_sweep(e, stream0, S ? *S : 0.0, mean, thr, check);
  HIPCHK(hipMemcpyAsync(e->sweep, e->h_sweep, sizeof(Sweep), hipMemcpyHostToDevice, e->stream));
  if (!S) {
    const int grid = launch_entropy(e);
    k_entropy_to_sweep<<<1, 256, 0, e->stream>>>(e->ent_part, grid,
                                                 usap_model_S(e->N, e->E, e->B), e->sweep);
    HIPCHK(hipGetLastError());
  }
  int32_t rebuilds = 0;
  if (int rc = run_sweep(e, batch, 0, nb, &rebuilds)) return rc;
  const Sweep& hs = *e->h_sweep;
  const int32_t last = hs.done - 1;  // the last executed batch
  e->last_beg = last * batch;
  e->last_end = std::min<int64_t>((int64_t)e->last_beg + batch, e->N);
  out->batches = hs.done;
  out->proposals = std::min<int64_t>((int64_t)hs.done * batch, e->N);
  if (e->uni) e->uni_pos += out->proposals;
  out->converged = hs.stop == 1;
  out->moves = (int64_t)hs.moves;
  out->S = hs.S;
  out->dS_sum = hs.dS_last;
  out->dS_hi = (int64_t)hs.cum_hi;
  out->dS_lo = (uint64_t)hs.cum_lo;
  out->dS_inf = hs.cum_inf ? -1 : 0;
  out->rebuilds = rebuilds;
  return USAP_OK;
}

}  // namespace usap

extern "C" {

int usap_nodal_sum_fx(usap_engine* e, int64_t* hi, uint64_t* lo, int32_t* inf) {
  if (int rc = check_engine(e)) return rc;
  const Sweep& hs = *e->h_sweep;
  if (hi) *hi = (int64_t)hs.cum_hi;
  if (lo) *lo = (uint64_t)hs.cum_lo;
  if (inf) *inf = hs.cum_inf ? -1 : 0;
  return USAP_OK;
}

int usap_nodal_detail(usap_engine* e, int32_t* s, int64_t* hi, uint64_t* lo, int32_t* inf,
                      int32_t* accept) {
  if (int rc = check_engine(e)) return rc;
  const int64_t n = e->last_end - e->last_beg;
  hipStream_t st = e->stream;
  if (n > 0) {
    if (s) HIPCHK(hipMemcpyAsync(s, e->prop_s, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    if (hi) HIPCHK(hipMemcpyAsync(hi, e->prop_hi, sizeof(int64_t) * n, hipMemcpyDeviceToHost, st));
    if (lo) HIPCHK(hipMemcpyAsync(lo, e->prop_lo, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, st));
    if (inf) HIPCHK(hipMemcpyAsync(inf, e->prop_inf, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    if (accept) HIPCHK(hipMemcpyAsync(accept, e->prop_acc, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  return USAP_OK;
}

int usap_profile(usap_engine* e, int32_t enable, usap_kernel_prof* out) {
  if (!e) return USAP_EINVAL;
  if (out) {
    for (int k = 0; k < USAP_NKERNELS; ++k) out[k] = e->kp[k];
    if (e->aq_acc && e->prof) {
      unsigned long long acc[3 * kAqSlots];
      HIPCHK(hipMemcpyAsync(acc, e->aq_acc, sizeof(acc), hipMemcpyDeviceToHost, e->stream));
      HIPCHK(hipStreamSynchronize(e->stream));
      for (int j = 0; j < 2 * kAqSlots; ++j) out[USAP_K_NODAL].model_bytes += (double)acc[j];
      for (int j = 0; j < kAqSlots; ++j) out[USAP_K_MERGE].model_bytes += (double)acc[2 * kAqSlots + j];
    }
  }
  if (enable == 1) {
    for (int k = 0; k < USAP_NKERNELS; ++k) e->kp[k] = usap_kernel_prof{};
    if (!e->prop_u) {
      if (int rc = dalloc(e, &e->prop_u, e->N)) return rc;
      if (int rc = dalloc(e, &e->aq_mark, e->N)) return rc;
      if (int rc = dalloc(e, &e->aq_acc, 3 * kAqSlots)) return rc;
      HIPCHK(hipMemsetAsync(e->aq_mark, 0xff, sizeof(int32_t) * e->N, e->stream));
      e->aq_stamp = 0;
    }
    HIPCHK(hipMemsetAsync(e->aq_acc, 0, 3 * kAqSlots * sizeof(unsigned long long), e->stream));
    e->prof = true;
  } else if (enable == 0) {
    e->prof = false;
  }
  return USAP_OK;
}

int usap_partition_log(usap_engine* e, int32_t cap, usap_phase* out, int32_t* n) {
  if (!e || !n) return USAP_EINVAL;
  const int32_t m = (int32_t)e->phase_log.size();
  for (int32_t i = 0; i < m && i < cap && out; ++i) out[i] = e->phase_log[i];
  *n = m;
  return USAP_OK;
}

int usap_snapshot(usap_engine* e, int32_t slot, double S) {
  if (int rc = check_engine(e)) return rc;
  if (slot < 0 || slot >= kNumSnap) return USAP_EINVAL;
  if (!e->snap_labels[slot])
    if (int rc = dalloc(e, &e->snap_labels[slot], e->N)) return rc;
  HIPCHK(hipMemcpyAsync(e->snap_labels[slot], e->labels, sizeof(int32_t) * e->N,
                        hipMemcpyDeviceToDevice, e->stream));
  e->snap_B[slot] = e->B;
  e->snap_S[slot] = S;
  return USAP_OK;
}

int usap_restore(usap_engine* e, int32_t slot) {
  if (int rc = check_engine(e)) return rc;
  if (slot < 0 || slot >= kNumSnap || !e->snap_labels[slot]) return USAP_EINVAL;
  HIPCHK(hipMemcpyAsync(e->labels, e->snap_labels[slot], sizeof(int32_t) * e->N,
                        hipMemcpyDeviceToDevice, e->stream));
  e->B = e->snap_B[slot];
  e->merge_B = -1;
  return rebuild_M(e);
}

}  // extern "C"
