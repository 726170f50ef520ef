// cac_ks_kernels.h -- device side of the dense Karger-Stein recursive contraction.
//
// One workgroup contracts one child of the recursion tree (indexed_adjacency_matrix /
// lazy_adjacency_matrix semantics, src/karger-stein/indexed_adjacency_matrix.hpp:137-191,
// lazy_adjacency_matrix.hpp:120-216):
//   - the parent matrix and its row sums are shared read-only by both children; a
//     child writes only the rows it merges into, into rows taken from a level-wide
//     pool (copy on write, the row's pool offset kept in LDS);
//   - capacities live in LDS with 64-entry block sums, so the row pick is a
//     two-level search by one wavefront;
//   - every thread owns the same entries of every row for the whole kernel,
//     interleaved (thread t of B owns x = k B + t, k < 8), so each load or store
//     instruction of a wavefront covers 512 contiguous bytes of a row; the picked
//     row stays in registers for the column pick and the merge, the last merged
//     row stays there too and is written to its pool row only when another
//     merged row takes its place while it is live, or when the child ends, and
//     a thread only ever reads back what it wrote itself;
//   - a PR pass (co_mincut.h:50-63) may follow; the kernel ends by writing the
//     column map and the list of live rows, and a separate many-workgroup kernel
//     compacts the live rows into the child's t x t matrix with its row sums and
//     row maxima.
// Leaves (<= 128 vertices) get a lower-bound certificate from those row sums and
// maxima; Stoer-Wagner runs only on the leaves it does not settle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cac_common.h"
#include "cac_wave.h"

namespace cac {

constexpr uint32_t kKsBase = 128;        // base_case_size (co_mincut.h:24)
constexpr uint32_t kKsMaxS = 8192;       // LDS holds caps, reps and row slots for s <= this
constexpr uint32_t kKsChunk = 8;         // row entries per thread (s <= 8 * block size), x = k * B + t
constexpr uint32_t kKsUnit = 64;         // pool allocation unit: 64 int64 (512 bytes)
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
constexpr uint32_t kLiveWork = 0x80000000u;  // live[] entry: row taken from the pool
constexpr uint32_t kKsBad = 0xFFFFFFFFu;       // k_out: selection out of range
constexpr uint32_t kKsOverflow = 0xFFFFFFFEu;  // k_out: row pool exhausted, rerun the level

struct KsChild {
  const int64_t* parent;  // s x s, row stride s
  const int64_t* prsum;   // parent row sums [s]
  int64_t* out;           // k x k compacted (k <= tgt), row stride k
  int64_t* orsum;         // [tgt] row sums of out
  int64_t* ormax;         // [tgt] row maxima of out
  uint32_t* colmap;       // [s] column -> output index
  uint32_t* live;         // [tgt] output row -> parent row, or kLiveWork | pool unit
  uint32_t s, tgt, trial, nid;
  uint32_t slot, seed;  // seed: Philox key word 0 of the child's trial
  int64_t bound;
};

struct KsPool {
  int64_t* base;                    // merged rows, kKsUnit-aligned
  unsigned long long* used_units;   // device bump counter
  unsigned long long cap_units;
};

// Compaction work item: rows [r0, r0 + nr) of child c (rows >= k are skipped).
struct KsCompactItem {
  uint32_t child, r0, nr, pad;
};

struct KsLeaf {
  const int64_t* mat;   // k x k, row stride k
  const int64_t* rsum;  // [k] row sums, or null
  const int64_t* rmax;  // [k] row maxima, or null
  uint32_t k, slot;
  int64_t bound;
};

// ---- wave / block reductions (wave primitives in cac_wave.h) --------------------
// Block sum over B threads; every thread gets the total.  lds: >= B/64 u64.
template <int B>
__device__ inline uint64_t block_sum_u64(uint64_t x, uint64_t* lds) {
  x = wave_sum<uint64_t>(x);
  if (B == 64) return x;
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = x;
  __syncthreads();
  uint64_t t = 0;
#pragma unroll
  for (int w = 0; w < B / 64; ++w) t += lds[w];
  __syncthreads();
  return t;
}
// Block sum without the trailing barrier: the caller reaches a barrier before
// lds is written again.
template <int B>
__device__ inline uint64_t block_sum_u64_open(uint64_t x, uint64_t* lds) {
  x = wave_sum<uint64_t>(x);
  if (B == 64) return x;
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = x;
  __syncthreads();
  uint64_t t = 0;
#pragma unroll
  for (int w = 0; w < B / 64; ++w) t += lds[w];
  return t;
}
template <int B>
__device__ inline uint32_t block_min_u32(uint32_t x, uint32_t* lds) {
  x = wave_min<uint32_t>(x);
  if (B == 64) return x;
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = x;
  __syncthreads();
  uint32_t t = 0xFFFFFFFFu;
#pragma unroll
  for (int w = 0; w < B / 64; ++w) t = min(t, lds[w]);
  __syncthreads();
  return t;
}

// First index i in [0, s) whose inclusive prefix of cap reaches r (1 <= r <= total),
// by one wavefront: 64-entry block sums first (lane l owns blocks 2l, 2l + 1),
// then the 64 entries of the chosen block.  kNoSlot when r exceeds the total.
__device__ inline uint32_t wave_select_two_level(const int64_t* cap, const uint64_t* cbs, uint32_t s, uint32_t nb,
                                                 uint64_t r) {
  const int lane = threadIdx.x & 63;
  const uint32_t b0 = 2 * lane, b1 = b0 + 1;
  const uint64_t c0 = b0 < nb ? cbs[b0] : 0, c1 = b1 < nb ? cbs[b1] : 0;
  const uint64_t incl = wave_incl_scan<uint64_t>(c0 + c1);
  const uint64_t hit = __ballot(incl >= r);
  if (!hit) return kNoSlot;
  const int owner = __ffsll((long long)hit) - 1;
  const uint64_t before = readlane_u64(incl - c0 - c1, owner);
  const uint64_t oc0 = readlane_u64(c0, owner);
  uint32_t blk = 2 * owner;
  uint64_t rr = r - before;
  if (rr > oc0) {
    rr -= oc0;
    blk += 1;
  }
  const uint32_t x = blk * 64 + lane;
  const uint64_t v = x < s ? (uint64_t)cap[x] : 0;
  const uint64_t inc2 = wave_incl_scan<uint64_t>(v);
  const uint64_t hit2 = __ballot(x < s && inc2 >= rr);
  if (!hit2) return kNoSlot;
  return blk * 64 + (uint32_t)(__ffsll((long long)hit2) - 1);
}

// First index (natural column order) of a row whose inclusive prefix reaches r
// (1 <= r), with the row held in registers interleaved: thread t owns entries
// x = k * B + t, k < cc, and holds 0 for x >= s.  Natural order is segment k
// by segment k, so every wave's segment totals go to LDS (one barrier); then
// each thread finds the segment where the prefix reaches r and its wave's
// offset inside it, and the one thread whose entry holds r publishes its index
// (a second barrier), with rep[index] in lds_idx[1]: it reads that entry before
// the barrier, so a merge that relabels representatives right after cannot race
// with it.  lds: kKsChunk * B/64 u64.  kNoSlot if the row total is below r.
// The caller reaches a barrier before lds or lds_idx is written again.
template <int B>
__device__ inline uint32_t block_select_row(const int64_t (&v)[kKsChunk], uint32_t cc, uint64_t r, uint64_t* lds,
                                            uint32_t* lds_idx, const uint32_t* rep) {
  constexpr int W = B / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (uint32_t k = 0; k < kKsChunk; ++k) {
    if (k < cc) {
      const uint64_t t = wave_sum<uint64_t>((uint64_t)v[k]);
      if (lane == 0) lds[k * W + wave] = t;
    }
  }
  __syncthreads();
  uint64_t before = 0, val = 0;
  uint32_t ks = kNoSlot;
#pragma unroll
  for (uint32_t k = 0; k < kKsChunk; ++k) {
    if (k < cc && ks == kNoSlot) {
      uint64_t tot = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) tot += lds[k * W + w];
      if (before + tot >= r) {
        ks = k;
        val = (uint64_t)v[k];
      } else {
        before += tot;
      }
    }
  }
  if (ks == kNoSlot) return kNoSlot;  // the same for every thread
  uint64_t wb = 0;
  for (int w = 0; w < wave; ++w) wb += lds[ks * W + w];
  const uint64_t inc = wave_incl_scan<uint64_t>(val);
  const uint64_t ex = before + wb + inc - val;
  if (val != 0 && ex < r && r <= ex + val) {  // exactly one thread
    lds_idx[0] = ks * B + threadIdx.x;
    lds_idx[1] = rep[ks * B + threadIdx.x];
  }
  __syncthreads();
  return *lds_idx;
}

}  // namespace cac
