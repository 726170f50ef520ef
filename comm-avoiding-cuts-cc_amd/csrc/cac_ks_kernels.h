// cac_ks_kernels.h -- device side of the dense Karger-Stein recursive contraction.
//
// One workgroup contracts one child of the recursion tree (indexed_adjacency_matrix /
// lazy_adjacency_matrix semantics, src/karger-stein/indexed_adjacency_matrix.hpp:137-191,
// lazy_adjacency_matrix.hpp:120-216):
//   - the parent matrix is shared read-only by both children; a child writes only
//     the rows it merges into (copy on write, tracked per row in LDS), so no copy
//     of the s x s parent is ever made;
//   - capacities (weighted degrees) and representatives live in LDS;
//   - each pick: row i proportional to capacity, column k proportional to row i,
//     j = rep[k], merge max(i,j) into min(i,j): rep relabel, row add with loop
//     zeroing, fused capacity reduction;
//   - an optional PR pass (co_mincut.h:50-63) and the rank compaction into the
//     child's t x t output with LDS int64 atomics.
// Leaves (<= 128 vertices) are solved by one wavefront each: a lower-bound
// certificate first, Stoer-Wagner in LDS when the certificate does not settle it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cac_common.h"

namespace cac {

constexpr int kKsBlock = 256;
constexpr int kKsWaves = kKsBlock / 64;
constexpr uint32_t kKsBase = 128;       // base_case_size (co_mincut.h:24)
constexpr uint32_t kKsMaxLdsS = 8192;   // LDS holds cap + rep + row flags for s <= this

struct KsChild {
  const int64_t* parent;  // s x s, row stride s
  int64_t* work;          // s x s rows written on first merge
  int64_t* out;           // compacted k x k (k <= tgt)
  uint32_t s, tgt, trial, nid;
  uint32_t slot;          // result slot of the trial
  uint32_t pad;
  int64_t bound;
};

struct KsLeaf {
  const int64_t* mat;  // k x k, row stride k
  uint32_t k, slot;
  int64_t bound;
};

// ---- wave / block reductions -------------------------------------------------
__device__ inline uint64_t wave_sum_u64(uint64_t x) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}
__device__ inline uint32_t wave_min_u32(uint32_t x) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) x = min(x, (uint32_t)__shfl_xor(x, d, 64));
  return x;
}

// Block sum; every thread gets the total.  lds: >= kKsWaves u64.
__device__ inline uint64_t block_sum_u64(uint64_t x, uint64_t* lds) {
  x = wave_sum_u64(x);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = x;
  __syncthreads();
  uint64_t t = 0;
#pragma unroll
  for (int w = 0; w < kKsWaves; ++w) t += lds[w];
  __syncthreads();
  return t;
}
__device__ inline uint32_t block_min_u32(uint32_t x, uint32_t* lds) {
  x = wave_min_u32(x);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = x;
  __syncthreads();
  uint32_t t = 0xFFFFFFFFu;
#pragma unroll
  for (int w = 0; w < kKsWaves; ++w) t = min(t, lds[w]);
  __syncthreads();
  return t;
}

// First index i in [0, s) whose inclusive prefix of get(.) reaches r (1 <= r <= total).
// Wave w owns a contiguous segment; segment totals locate the wave, which then
// scans its segment 64 elements at a time.  lds_u64: >= kKsWaves; lds_idx: 1 u32.
template <class Get>
__device__ inline uint32_t block_select(Get get, uint32_t s, uint64_t r, uint64_t* lds_u64, uint32_t* lds_idx) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) *lds_idx = 0xFFFFFFFFu;  // stays when r exceeds the total
  const uint32_t seg = ((s + kKsWaves - 1) / kKsWaves + 63) & ~63u;
  const uint32_t lo = min(s, (uint32_t)wave * seg), hi = min(s, lo + seg);
  uint64_t part = 0;
  for (uint32_t x = lo + lane; x < hi; x += 64) part += (uint64_t)get(x);
  part = wave_sum_u64(part);
  if (lane == 0) lds_u64[wave] = part;
  __syncthreads();
  uint64_t before = 0;
  int owner = -1;
  for (int w = 0; w < kKsWaves; ++w) {
    uint64_t v = lds_u64[w];
    if (owner < 0 && before + v >= r) owner = w;
    else if (owner < 0) before += v;
  }
  if (wave == owner) {
    uint64_t carry = before;
    uint32_t found = 0xFFFFFFFFu;
    for (uint32_t base = lo; base < hi; base += 64) {
      uint32_t x = base + lane;
      uint64_t v = x < hi ? (uint64_t)get(x) : 0;
      uint64_t inc = v;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        uint64_t y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
      }
      uint64_t hit = __ballot(x < hi && carry + inc >= r);
      if (hit) {
        found = base + (uint32_t)(__ffsll((long long)hit) - 1);
        break;
      }
      carry += __shfl(inc, 63, 64);
    }
    if (lane == 0) *lds_idx = found;
  }
  __syncthreads();
  uint32_t res = *lds_idx;
  __syncthreads();
  return res;
}

// Same contract as block_select for a row in global memory: thread t owns the
// contiguous chunk [t*c, (t+1)*c), c = ceil(s / kKsBlock) <= kKsMaxChunk, and keeps
// it in registers, so one round of loads, one block scan and an in-register scan
// by the owning thread replace the serial wave walk.  lds_u64: >= kKsWaves + 1.
constexpr uint32_t kKsMaxChunk = kKsMaxLdsS / kKsBlock;
__device__ inline uint32_t block_select_row(const int64_t* row, uint32_t s, uint64_t r, uint64_t* lds_u64,
                                            uint32_t* lds_idx) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t c = (s + kKsBlock - 1) / kKsBlock;
  const uint32_t lo = threadIdx.x * c;
  if (threadIdx.x == 0) *lds_idx = 0xFFFFFFFFu;
  uint64_t v[kKsMaxChunk];
  uint64_t sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < kKsMaxChunk; ++k) {
    const uint32_t x = lo + k;
    v[k] = (k < c && x < s) ? (uint64_t)row[x] : 0ull;
    sum += v[k];
  }
  // block exclusive scan of the chunk sums
  uint64_t inc = sum;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t y = __shfl_up(inc, d, 64);
    if (lane >= d) inc += y;
  }
  if (lane == 63) lds_u64[wave] = inc;
  __syncthreads();
  uint64_t before = 0;
  for (int w = 0; w < wave; ++w) before += lds_u64[w];
  const uint64_t ex = before + inc - sum;
  if (sum != 0 && ex < r && r <= ex + sum) {
    uint64_t acc = ex;
    uint32_t found = 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t k = 0; k < kKsMaxChunk; ++k) {
      acc += v[k];
      if (found == 0xFFFFFFFFu && v[k] != 0 && acc >= r) found = lo + k;
    }
    *lds_idx = found;
  }
  __syncthreads();
  const uint32_t res = *lds_idx;
  __syncthreads();
  return res;
}

}  // namespace cac
