// usap_common.hpp -- arithmetic and random-stream helpers shared by the host
// driver and the HIP kernels of the engine.
//
// lambda(x) = llrint(log(x) * 2^48) is an int64 table built once on the host;
// an entropy term T(m,x,y) = m*(lambda(m)-lambda(x)-lambda(y)) is an exact
// int128, so every summation order and every algebraic regrouping of the
// reference's dS sums (graph_parallel.hpp:867-987, :1096-1215) gives the same
// bits on the CPU and on the GPU.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define USAP_HD __host__ __device__ __forceinline__
#else
#define USAP_HD inline
#endif

namespace usap {

typedef __int128 i128;
typedef unsigned __int128 u128;

constexpr int kLambdaFrac = 48;
constexpr uint64_t kTagMerge = 1;
constexpr uint64_t kTagNodal = 2;

// Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11): the paper's round
// multipliers and key increments.
USAP_HD void philox4x64(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3, uint64_t k0,
                        uint64_t k1, uint64_t out[4]) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int round = 0; round < 10; ++round) {
    if (round > 0) {
      k0 += 0x9E3779B97F4A7C15ull;
      k1 += 0xBB67AE8584CAA73Bull;
    }
    const u128 p0 = (u128)0xD2E7470EE14C6C93ull * c0;
    const u128 p1 = (u128)0xCA5A826395121157ull * c2;
    const uint64_t n0 = (uint64_t)(p1 >> 64) ^ c1 ^ k0;
    const uint64_t n2 = (uint64_t)(p0 >> 64) ^ c3 ^ k1;
    c1 = (uint64_t)p1;
    c3 = (uint64_t)p0;
    c0 = n0;
    c2 = n2;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// index in [0, W): floor(x * W / 2^64)
USAP_HD int64_t idx_of(uint64_t x, int64_t W) {
  return (int64_t)(((u128)x * (u128)(uint64_t)W) >> 64);
}

// accept-uniform branch of graph_parallel.hpp:816 and :1034,
// x <= B / (d_a[u] + B) with x = (x1 >> 11) * 2^-53, decided exactly.
USAP_HD bool uniform_branch(uint64_t x1, int64_t da_u, int64_t B) {
  return (u128)(x1 >> 11) * (u128)(uint64_t)(da_u + B) <= ((u128)(uint64_t)B << 53);
}

// uniform block different from r, graph_parallel.hpp:817-821
USAP_HD int32_t uniform_not_r(uint64_t x2, int32_t r, int32_t B) {
  int32_t s = (int32_t)idx_of(x2, B);
  if (s == r) s++;
  if (s == B) s = 0;
  return s;
}

}  // namespace usap
