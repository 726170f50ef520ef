// cac_ks.hip -- dense Karger-Stein on gfx950: breadth-first recursive contraction
// over all trials at once, a many-workgroup compaction per level, and
// certificate-first leaves.
//
// Replaces comincut::minimum_cut_trials_start (dense, src/karger-stein/co_mincut.h:370-390),
// minimum_cut_trials (:315-342), recursive_contraction_inner (:176-223),
// contract_graph (:38-66) and the base case deterministic_cut_pr
// (co_mincut_base_case.hpp:164-185).  Sizes follow
// target_number_of_vertices_for_contraction (co_mincut.cpp:18-20).
//
// Per level of the recursion tree:
//   1. k_ks_contract<B>: one workgroup per child (B = 64, 256 or 1024 threads by
//      size) draws s - t' picks, merges rows into pool rows, runs the PR pass and
//      writes the column map and the live-row list;
//   2. k_ks_compact: workgroups of a few output rows each sum live rows by column
//      class into the child's t' x t' matrix, with row sums and row maxima;
//   3. children of at most 128 vertices are leaves: k_ks_cert settles most of them
//      from row sums and maxima; k_ks_sw runs Stoer-Wagner on the rest.
// Device memory for a level comes from two grow-only buffers of the context used
// alternately (the parents live in one, the children in the other).
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "cac_internal.h"
#include "cac_ks.h"
#include "cac_ks_kernels.h"

namespace cac {
namespace {

// ---- leaves ------------------------------------------------------------------
// One wavefront per leaf: lane l owns vertices l and l + 64.
constexpr int kLeafMax = 128;
// Row stride of a leaf matrix in LDS: one entry more than a row, so a column
// (lane x reading or writing entry (x, c), as a merge and the PR scan do) spans
// the banks instead of hitting one: with 128 int64 per row every row starts on
// the same bank and a column access is a 64-way bank conflict.
constexpr int kLeafStride = kLeafMax + 1;

__device__ inline void wave_argmax(int64_t& key, uint32_t& idx) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    int64_t k2 = __shfl_xor(key, d, 64);
    uint32_t i2 = __shfl_xor(idx, d, 64);
    if (k2 > key || (k2 == key && i2 < idx)) {
      key = k2;
      idx = i2;
    }
  }
}

__device__ inline int64_t wave_sum_i64(int64_t x) { return (int64_t)wave_sum<uint64_t>((uint64_t)x); }

// d - (q-1) M, saturating to a large negative value instead of overflowing
// (a negative term only weakens the bound).
__device__ inline int64_t leaf_f(int64_t d, int64_t M, uint32_t q) {
  const int64_t k = (int64_t)q - 1;
  if (M != 0 && k > INT64_MAX / 256 / M) return INT64_MIN / 256;
  return d - k * M;
}

// Lower bound on the minimum cut of a leaf: for a side S with |S| = q <= V/2,
// cut(S) = sum over u in S of (d(u) - w(u, S \ u)) >= sum of f_q(u) over S, with
// f_q(u) = d(u) - (q-1) M(u) and M(u) the heaviest edge at u; so the min over q
// of the sum of the q smallest f_q bounds every cut from below.
__device__ int64_t leaf_lower_bound(const int64_t d[2], const int64_t M[2], const bool alive[2], uint32_t V,
                                    int64_t stop_at) {
  const int lane = threadIdx.x & 63;
  int64_t best = INT64_MAX;
  for (uint32_t q = 1; q <= V / 2; ++q) {
    int64_t f0 = alive[0] ? leaf_f(d[0], M[0], q) : INT64_MAX;
    int64_t f1 = alive[1] ? leaf_f(d[1], M[1], q) : INT64_MAX;
    // bitonic sort of 128 keys: element e = lane + 64 * r lives in register r
#pragma unroll
    for (int k = 2; k <= 128; k <<= 1) {
#pragma unroll
      for (int j = k >> 1; j > 0; j >>= 1) {
        if (j == 64) {
          // partner of lane's element 0 is its element 1
          bool up = (lane & k) == 0;  // k == 128 -> up for all
          if ((f0 > f1) == up) {
            int64_t t = f0;
            f0 = f1;
            f1 = t;
          }
        } else {
          int64_t p0 = __shfl_xor(f0, j, 64), p1 = __shfl_xor(f1, j, 64);
          bool lower = (lane & j) == 0;
          bool up0 = ((lane) & k) == 0, up1 = ((lane + 64) & k) == 0;
          int64_t mn0 = f0 < p0 ? f0 : p0, mx0 = f0 < p0 ? p0 : f0;
          int64_t mn1 = f1 < p1 ? f1 : p1, mx1 = f1 < p1 ? p1 : f1;
          f0 = (lower == up0) ? mn0 : mx0;
          f1 = (lower == up1) ? mn1 : mx1;
        }
      }
    }
    // sum of the q smallest: elements 0..q-1 (element e in register e / 64 of lane e % 64)
    int64_t part = ((uint32_t)lane < q ? f0 : 0) + ((uint32_t)(lane + 64) < q ? f1 : 0);
    const int64_t lb = wave_sum_i64(part);
    if (lb < best) best = lb;
    if (best < stop_at) break;
  }
  return best;
}

// Certificate first: a leaf whose lower bound reaches the trial's bound
// contributes the bound; the others are queued for Stoer-Wagner.
__global__ void __launch_bounds__(64) k_ks_cert(const KsLeaf* __restrict__ leaves, int64_t* __restrict__ results,
                                                uint32_t* __restrict__ sw_list, uint32_t* __restrict__ sw_count) {
  const KsLeaf lf = leaves[blockIdx.x];
  const uint32_t V = lf.k;
  const int lane = threadIdx.x;
  const bool alive[2] = {(uint32_t)lane < V, (uint32_t)lane + 64 < V};
  int64_t d[2] = {0, 0}, M[2] = {0, 0};
  for (int h = 0; h < 2; ++h) {
    const uint32_t u = lane + 64 * h;
    if (!alive[h]) continue;
    if (lf.rsum && lf.rmax) {
      d[h] = lf.rsum[u];
      M[h] = lf.rmax[u];
    } else {
      int64_t sm = 0, mx = 0;
      for (uint32_t x = 0; x < V; ++x) {
        const int64_t v = lf.mat[(size_t)u * V + x];
        sm += v;
        mx = v > mx ? v : mx;
      }
      d[h] = sm;
      M[h] = mx;
    }
  }
  const int64_t lb = V >= 2 ? leaf_lower_bound(d, M, alive, V, lf.bound) : 0;
  // leaves whose total weight is below 2^32 (every entry and every merged
  // entry fits 32 bits) go to the front of the list, the others to the back
  const uint64_t total = wave_sum<uint64_t>((uint64_t)(d[0] + d[1]));
  if (lane == 0) {
    if (lb >= lf.bound) atomicMin((long long*)&results[lf.slot], (long long)lf.bound);
    else if (total < (1ull << 32)) sw_list[atomicAdd(&sw_count[0], 1u)] = blockIdx.x;
    else sw_list[gridDim.x - 1 - atomicAdd(&sw_count[1], 1u)] = blockIdx.x;
  }
}

// Merge vertex `last` into `prev` in the LDS matrix (row and column); lanes own
// columns lane and lane + 64.
template <typename T>
__device__ inline void leaf_merge(T* A, uint32_t V, uint32_t prev, uint32_t last, bool (&alive)[2]) {
  const int lane = threadIdx.x;
  for (int h = 0; h < 2; ++h) {
    const uint32_t x = lane + 64 * h;
    if (x < V && x != prev && x != last) {
      const T v = A[prev * kLeafStride + x] + A[last * kLeafStride + x];
      A[prev * kLeafStride + x] = v;
      A[x * kLeafStride + prev] = v;
    }
  }
  __syncthreads();
  if (lane == 0) {
    A[prev * kLeafStride + last] = 0;
    A[last * kLeafStride + prev] = 0;
  }
  if ((uint32_t)lane == last) alive[0] = false;
  if ((uint32_t)lane + 64 == last) alive[1] = false;
  __syncthreads();
}

// The base case of co_mincut_base_case.hpp:164-185 on a leaf in LDS: PR-1 pass
// at the bound (contracting an edge of weight >= bound keeps every cut below the
// bound, so min(bound, min cut) is unchanged), then Stoer-Wagner (J. ACM 44(4),
// 1997): maximum adjacency order, cut of the phase, merge of the last two
// vertices.  Result min(cut, bound).  T: the LDS entry type, uint32_t for a
// leaf of total weight below 2^32 (66 KiB, two leaves per CU) and int64_t
// otherwise (132 KiB, one per CU).
template <typename T>
__global__ void __launch_bounds__(64) k_ks_sw(const KsLeaf* __restrict__ leaves, const uint32_t* __restrict__ sw_list,
                                              int64_t* __restrict__ results) {
  __shared__ T A[kLeafMax * kLeafStride];
  const KsLeaf lf = leaves[sw_list[blockIdx.x]];
  const uint32_t V = lf.k;
  const int lane = threadIdx.x;
  int64_t total = 0;  // sum of all entries: no key of the maximum adjacency order exceeds it
  for (uint32_t r = 0; r < V; ++r)
    for (uint32_t x = lane; x < V; x += 64) {
      const int64_t a = lf.mat[(size_t)r * V + x];
      A[r * kLeafStride + x] = (T)a;
      total += a;
    }
  total = wave_sum_i64(total);
  // keys below 2^55 pack with their vertex as (key << 8) | (255 - vertex), so the
  // largest key, then the smallest vertex, is one unsigned wave maximum
  const bool packed = total < (int64_t)1 << 55;
  __syncthreads();
  bool alive[2] = {(uint32_t)lane < V, (uint32_t)(lane + 64) < V};
  int64_t best = INT64_MAX;
  uint32_t nalive = V;
  // PR-1: repeatedly contract an edge of weight >= bound (the lowest live row
  // holding one, its lowest live column); merged-away columns hold stale values
  // and are skipped.  col[h] is the first such column of row lane + 64 h.  A
  // merge of last into prev (prev < last) changes only row prev and column
  // prev among the live entries, and only upwards, so a row other than prev
  // keeps its first column unless column prev now qualifies before it; a row
  // whose first column was last finds prev there (it took last's weight), and
  // row prev alone is scanned again.
  auto first_col = [&](uint32_t u) {
    const uint64_t live_lo = __ballot(alive[0]), live_hi = __ballot(alive[1]);
    uint32_t c = kNoSlot;
    for (uint32_t x = 0; x < V; ++x) {
      const bool live_x = ((x < 64 ? live_lo : live_hi) >> (x & 63)) & 1ull;
      if (live_x && x != u && (int64_t)A[u * kLeafStride + x] >= lf.bound) {
        c = x;
        break;
      }
    }
    return c;
  };
  uint32_t col[2] = {kNoSlot, kNoSlot};
  {
    const uint32_t c0 = first_col((uint32_t)lane), c1 = first_col((uint32_t)lane + 64);
    col[0] = alive[0] ? c0 : kNoSlot;
    col[1] = alive[1] ? c1 : kNoSlot;
  }
  while (nalive > 1) {
    const uint64_t lo_mask = __ballot(col[0] != kNoSlot), hi_mask = __ballot(col[1] != kNoSlot);
    if (!lo_mask && !hi_mask) break;
    const uint32_t u = lo_mask ? (uint32_t)(__ffsll((long long)lo_mask) - 1)
                               : 64u + (uint32_t)(__ffsll((long long)hi_mask) - 1);
    const uint32_t x = __shfl(u < 64 ? col[0] : col[1], (int)(u & 63), 64);
    const uint32_t prev = u < x ? u : x, last = u < x ? x : u;
    leaf_merge(A, V, prev, last, alive);
    --nalive;
    // row prev again, scanned by the whole wavefront: lane l tests columns l
    // and l + 64, and the first set bit of the ballots is the first column
    uint32_t again;
    {
      const bool q0 = alive[0] && (uint32_t)lane != prev && (int64_t)A[prev * kLeafStride + lane] >= lf.bound;
      const bool q1 =
          alive[1] && (uint32_t)lane + 64 != prev && (int64_t)A[prev * kLeafStride + lane + 64] >= lf.bound;
      const uint64_t b0 = __ballot(q0), b1 = __ballot(q1);
      again = b0 ? (uint32_t)(__ffsll((long long)b0) - 1) : b1 ? 64u + (uint32_t)(__ffsll((long long)b1) - 1) : kNoSlot;
    }
    for (int h = 0; h < 2; ++h) {
      const uint32_t r = lane + 64 * h;
      if (!alive[h]) col[h] = kNoSlot;
      else if (r == prev) col[h] = again;
      else if (col[h] == last) col[h] = prev;
      else if (prev < col[h] && (int64_t)A[r * kLeafStride + prev] >= lf.bound) col[h] = prev;
    }
  }
  if (lf.bound <= 0) nalive = 1;  // a non-positive bound is already the answer
  while (nalive > 1) {
    int64_t key[2] = {0, 0};
    bool added[2] = {!alive[0], !alive[1]};
    // start from the smallest live vertex
    uint64_t live_lo = __ballot(alive[0]), live_hi = __ballot(alive[1]);
    uint32_t start =
        live_lo ? (uint32_t)(__ffsll((long long)live_lo) - 1) : 64u + (uint32_t)(__ffsll((long long)live_hi) - 1);
    uint32_t prev = start, last = start;
    if ((uint32_t)lane == start) added[0] = true;
    if ((uint32_t)lane + 64 == start) added[1] = true;
    for (int h = 0; h < 2; ++h) key[h] += (int64_t)A[start * kLeafStride + lane + 64 * h];
    int64_t cut = 0;
    for (uint32_t step = 1; step < nalive; ++step) {
      int64_t bk = INT64_MIN;
      uint32_t bi = 0xFFFFFFFFu;
      if (packed) {
        const uint64_t p0 = added[0] ? 0 : ((uint64_t)key[0] << 8) | (255u - (uint32_t)lane);
        const uint64_t p1 = added[1] ? 0 : ((uint64_t)key[1] << 8) | (255u - (uint32_t)lane - 64u);
        const uint64_t b = wave_max<uint64_t>(p0 > p1 ? p0 : p1);
        if (b != 0) {
          bk = (int64_t)(b >> 8);
          bi = 255u - (uint32_t)(b & 255u);
        }
      } else {
        for (int h = 0; h < 2; ++h) {
          if (!added[h] && (key[h] > bk || (key[h] == bk && (uint32_t)lane + 64 * h < bi))) {
            bk = key[h];
            bi = lane + 64 * h;
          }
        }
        wave_argmax(bk, bi);
      }
      if (bi >= V) break;  // no candidate: cannot happen for step < nalive
      prev = last;
      last = bi;
      if ((uint32_t)lane == bi) added[0] = true;
      if ((uint32_t)lane + 64 == bi) added[1] = true;
      if (step + 1 == nalive) {
        cut = bk;
        break;
      }
      for (int h = 0; h < 2; ++h)
        if (!added[h]) key[h] += (int64_t)A[bi * kLeafStride + lane + 64 * h];
    }
    best = cut < best ? cut : best;
    leaf_merge(A, V, prev, last, alive);
    --nalive;
  }
  best = best < lf.bound ? best : lf.bound;
  if (lane == 0) atomicMin((long long*)&results[lf.slot], (long long)best);
}

// ---- contraction -------------------------------------------------------------
// LDS layout for a child of s vertices: cap[s] i64 | cbs[nb] u64 | rep[s] u32 |
// wslot[s] u32 | red[B/64] u64 | misc[8] u32.
__host__ __device__ inline size_t contract_lds_bytes(uint32_t s, int B) {
  const size_t nb = (s + 63) / 64;
  return 8 * (size_t)s + 8 * nb + 8 * (size_t)s + 8 * (size_t)(kKsChunk + 1) * (B / 64) + 32;
}

struct KsLds {
  int64_t* cap;
  uint64_t* cbs;
  uint32_t* rep;
  uint32_t* wslot;
  uint64_t* red;
  uint32_t* misc;  // [0] i  [1] k  [2] rep[k]  [5] spare pool row
};

__device__ inline const int64_t* ks_row(const KsChild& c, const KsPool& pool, uint32_t slot, uint32_t r) {
  return slot == kNoSlot ? c.parent + (size_t)r * c.s : pool.base + (size_t)slot * kKsUnit;
}

// This thread's entries x = k * B + t of a row (0 past s).  The row pointer is
// generic (it comes from a struct in memory); as a global-address-space
// pointer the loads are global loads, which unlike flat loads do not count
// against the LDS wait counter, so the block's LDS waits between picks do not
// also wait for row loads in flight.
template <int B>
__device__ inline void load_row(const int64_t* row, uint32_t cc, uint32_t s, int64_t (&v)[kKsChunk]) {
  typedef const __attribute__((address_space(1))) int64_t g_i64;
  g_i64* g = (g_i64*)row;
#pragma unroll
  for (uint32_t k = 0; k < kKsChunk; ++k) {
    const uint32_t x = k * B + threadIdx.x;
    v[k] = (k < cc && x < s) ? g[x] : 0;
  }
}

template <int B>
__device__ inline void store_row(int64_t* row, uint32_t cc, uint32_t s, const int64_t (&v)[kKsChunk]) {
#pragma unroll
  for (uint32_t k = 0; k < kKsChunk; ++k) {
    const uint32_t x = k * B + threadIdx.x;
    if (k < cc && x < s) row[x] = v[k];
  }
}

// Merge row a into row b (a dies, b keeps its index).  The calling threads hold
// their chunk of row `held` (a or b) in v, and in u the last merged row, row
// `cached` (kNoSlot before the first merge), which is the other row when
// u_has_other.  That row is written back lazily: its pool row is stale until it
// leaves the registers.  Otherwise the other row is loaded into u here, after
// row `cached` is written to its pool row if it is still live and not a or b
// (merged into b's new row or dead, it needs no write).  On return u holds the
// merged row (0 past s), not yet written.  Its pool row is b's, else a's (a
// dies), else the child's spare pool row (misc[5]), which thread 0 then
// replaces with a fresh one.  Every thread derives that destination from the
// same LDS words; thread 0 changes them only after the block sum's barrier,
// which every thread reaches after its reads, and the merge ends with one more
// barrier.  Each thread relabels only the representatives of the columns it
// owns.  Returns the contracted weight (cap[a] + cap[b] - cap_new) / 2 in w;
// false when the pool is exhausted.  Called by the whole block.
template <int B>
__device__ bool ks_merge(const KsChild& c, const KsPool& pool, const KsLds& L, uint32_t a, uint32_t b, uint32_t held,
                         const int64_t (&v)[kKsChunk], int64_t (&u)[kKsChunk], uint32_t cached, bool u_has_other,
                         uint32_t cc, int64_t& w) {
  const uint32_t s = c.s;
  const uint32_t other = held == a ? b : a;
  const uint32_t slot_a = L.wslot[a], slot_b = L.wslot[b];
  const bool spare = slot_b == kNoSlot && slot_a == kNoSlot;
  const uint32_t db = slot_b != kNoSlot ? slot_b : slot_a != kNoSlot ? slot_a : L.misc[5];
  if (db == kNoSlot) return false;  // only a spare can be missing: the pool is exhausted
  if (!u_has_other) {
    if (cached != kNoSlot && cached != a && cached != b)  // a merged row has a pool row
      store_row<B>(pool.base + (size_t)L.wslot[cached] * kKsUnit, cc, s, u);
    load_row<B>(ks_row(c, pool, other == a ? slot_a : slot_b, other), cc, s, u);
  }
  uint64_t sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < kKsChunk; ++k) {
    const uint32_t x = k * B + threadIdx.x;
    if (k < cc && x < s) {
      int64_t t = v[k] + u[k];
      const uint32_t rx = L.rep[x];
      if (rx == a || rx == b) t = 0;  // loop of the merged vertex
      if (rx == a) L.rep[x] = b;
      u[k] = t;
      sum += (uint64_t)t;
    }
  }
  const int64_t ca = L.cap[a], cb = L.cap[b];
  const int64_t newcap = (int64_t)block_sum_u64_open<B>(sum, L.red);
  w = (ca + cb - newcap) / 2;
  if (B == 64) __syncthreads();  // every lane has read the slots and capacities
  if (threadIdx.x == 0) {
    if (slot_b == kNoSlot) {
      L.wslot[b] = db;
      L.wslot[a] = kNoSlot;
    }
    if (spare) {
      const unsigned long long units = (s + kKsUnit - 1) / kKsUnit;
      const unsigned long long off = atomicAdd(pool.used_units, units);
      L.misc[5] = off + units <= pool.cap_units ? (uint32_t)off : kNoSlot;
    }
    L.cbs[a / 64] -= (uint64_t)ca;
    L.cap[a] = 0;
    L.cbs[b / 64] += (uint64_t)(newcap - cb);
    L.cap[b] = newcap;
  }
  __syncthreads();
  return true;
}

// Below 1024 threads, at most 96 VGPRs (5 waves per SIMD): room for the two
// rows a thread keeps (the picked row and the last merged row) without
// spilling; at 80 VGPRs (6 waves) the kernel spills and is slower.
template <int B>
__global__ void __launch_bounds__(B, B >= 1024 ? 1 : 5)
    k_ks_contract(const KsChild* __restrict__ children, KsPool pool, uint32_t* __restrict__ k_out,
                  int64_t* __restrict__ results) {
  extern __shared__ __align__(16) unsigned char smem[];
  const KsChild c = children[blockIdx.x];
  const uint32_t s = c.s, nb = (s + 63) / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  KsLds L;
  L.cap = (int64_t*)smem;
  L.cbs = (uint64_t*)(smem + 8 * (size_t)s);
  L.rep = (uint32_t*)(smem + 8 * (size_t)s + 8 * nb);
  L.wslot = L.rep + s;
  L.red = (uint64_t*)(smem + 16 * (size_t)s + 8 * nb);
  L.misc = (uint32_t*)(L.red + (kKsChunk + 1) * (B / 64));

  for (uint32_t x = threadIdx.x; x < s; x += B) {
    L.cap[x] = c.prsum[x];
    L.rep[x] = x;
    L.wslot[x] = kNoSlot;
  }
  if (threadIdx.x == 0) {  // the spare pool row of the first merge of two original rows
    const unsigned long long units = (s + kKsUnit - 1) / kKsUnit;
    const unsigned long long off = atomicAdd(pool.used_units, units);
    L.misc[5] = off + units <= pool.cap_units ? (uint32_t)off : kNoSlot;
  }
  __syncthreads();
  for (uint32_t blk = wave; blk < nb; blk += B / 64) {
    const uint32_t x = blk * 64 + lane;
    const uint64_t t = wave_sum<uint64_t>(x < s ? (uint64_t)L.cap[x] : 0);
    if (lane == 0) L.cbs[blk] = t;
  }
  __syncthreads();
  uint64_t W = 0;
  for (uint32_t blk = 0; blk < nb; ++blk) W += L.cbs[blk];

  const uint32_t cc = (s + B - 1) / B;  // entries per thread, <= kKsChunk
  const uint32_t lo = threadIdx.x * cc;
  uint32_t nv = s, pr1 = 0;
  int status = 0;  // 1 disconnected, 2 bad selection, 3 pool exhausted
  // v: the row being picked or checked; u: the last merged row, whose index is
  // cb (most picks touch it: the heavy super-vertex is picked or is the picked
  // row's partner, and the PR pass checks it next after merging into it), so
  // it is not loaded again
  int64_t v[kKsChunk], u[kKsChunk];
  uint32_t cb = kNoSlot;
  for (uint32_t p = 0; p + c.tgt < s; ++p) {
    if (W == 0) {
      status = 1;
      break;
    }
    const u32x4 rnd = philox4x32(c.trial, c.nid, p, 0u, c.seed, kDomKs);
    if (wave == 0) {
      const uint32_t i = wave_select_two_level(L.cap, L.cbs, s, nb, uniform1(lo64(rnd), W));
      if (lane == 0) L.misc[0] = i;
    }
    __syncthreads();
    const uint32_t i = L.misc[0];
    if (i >= s) {
      status = 2;
      break;
    }
    if (i == cb) {
#pragma unroll
      for (uint32_t q = 0; q < kKsChunk; ++q) v[q] = u[q];
    } else {
      load_row<B>(ks_row(c, pool, L.wslot[i], i), cc, s, v);
    }
    const uint32_t k =
        block_select_row<B>(v, cc, uniform1(hi64(rnd), (uint64_t)L.cap[i]), L.red, L.misc + 1, L.rep);
    if (k >= s) {
      status = 2;
      break;
    }
    const uint32_t j = L.misc[2];  // rep[k], read by the selecting thread before the merge relabels
    const uint32_t a = i > j ? i : j, b = i > j ? j : i;
    int64_t w = 0;
    if (!ks_merge<B>(c, pool, L, a, b, i, v, u, cb, j == cb, cc, w)) {
      status = 3;
      break;
    }
    cb = b;
    W -= 2 * (uint64_t)w;
    --nv;
    if (w >= c.bound) ++pr1;
  }
  // PR pass (indexed_adjacency_matrix::pr_pass): contract every edge of weight >= bound
  if (status == 0 && pr1 > 0 && c.bound > 0) {
    for (uint32_t i = 0; i < s && nv > 2; ++i) {
      if (L.rep[i] != i || L.cap[i] < c.bound) continue;
      if (i == cb) {
#pragma unroll
        for (uint32_t q = 0; q < kKsChunk; ++q) v[q] = u[q];
      } else {
        load_row<B>(ks_row(c, pool, L.wslot[i], i), cc, s, v);
      }
      uint32_t first = kNoSlot;
#pragma unroll
      for (uint32_t k = 0; k < kKsChunk; ++k)
        if (first == kNoSlot && k < cc && k * B + threadIdx.x < s && v[k] >= c.bound) first = k * B + threadIdx.x;
      first = block_min_u32<B>(first, (uint32_t*)L.red);
      if (first == kNoSlot) continue;
      // rep[first] is not i: row i's entries of its own class are 0, below the
      // bound, so this merge's relabelling (of class i) leaves it unchanged
      int64_t w = 0;
      const uint32_t tgt = L.rep[first];
      if (!ks_merge<B>(c, pool, L, i, tgt, i, v, u, cb, tgt == cb, cc, w)) {
        status = 3;
        break;
      }
      cb = tgt;
      --nv;
    }
  }
  if (status == 0 && cb != kNoSlot) store_row<B>(pool.base + (size_t)L.wslot[cb] * kKsUnit, cc, s, u);
  if (status != 0) {
    if (threadIdx.x == 0) {
      k_out[blockIdx.x] = status == 1 ? 0u : status == 2 ? kKsBad : kKsOverflow;
      if (status == 1) atomicMin((long long*)&results[c.slot], 0ll);
    }
    return;
  }
  // ranks of the live rows (natural order), live-row list, column map
  __syncthreads();
  uint32_t* rank = (uint32_t*)L.cap;  // capacities are no longer needed
  uint32_t cnt = 0;
#pragma unroll
  for (uint32_t k = 0; k < kKsChunk; ++k)
    if (k < cc && lo + k < s && L.rep[lo + k] == lo + k) ++cnt;
  uint32_t inc = cnt;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(inc, d, 64);
    if (lane >= d) inc += y;
  }
  if (lane == 63) L.red[wave] = inc;
  __syncthreads();
  uint32_t ex = inc - cnt, total = 0;
  for (int w2 = 0; w2 < B / 64; ++w2) {
    if (w2 < wave) ex += (uint32_t)L.red[w2];
    total += (uint32_t)L.red[w2];
  }
#pragma unroll
  for (uint32_t k = 0; k < kKsChunk; ++k) {
    const uint32_t x = lo + k;
    if (k < cc && x < s && L.rep[x] == x) {
      rank[x] = ex;
      const uint32_t sl = L.wslot[x];
      c.live[ex] = sl == kNoSlot ? x : (kLiveWork | sl);
      ++ex;
    }
  }
  __syncthreads();
  for (uint32_t y = threadIdx.x; y < s; y += B) c.colmap[y] = rank[L.rep[y]];
  if (threadIdx.x == 0) k_out[blockIdx.x] = total;
}

// Rows per compaction workgroup for a child of t' output vertices (the LDS
// accumulator holds rows x t' int64, at most 32 KiB unless t' alone is larger).
constexpr uint32_t kCompactUnroll = 8;  // row elements per thread in flight in k_ks_compact

// Threads per compaction workgroup for children of tgt output vertices: the
// context's CAC_TUNE_KS_COMPACT_BLOCK setting for tgt's class, or 512 (on
// MI355X 17% less compaction time than 256 on the n = 8192 complete graph,
// and 1024 was slower than either).
inline int compact_block(const cac_ctx* ctx, uint32_t tgt) {
  const int cls = tgt <= 512 ? 0 : tgt <= 1024 ? 1 : tgt <= 2048 ? 2 : tgt <= 4096 ? 3 : 4;
  const int v = ctx->ks_compact_block[cls];
  return v ? v : 512;
}

__host__ __device__ inline uint32_t compact_rows(uint32_t tgt) {
  const uint32_t r = 4096 / (tgt ? tgt : 1);
  return r < 1 ? 1 : (r > 64 ? 64 : r);
}

// One workgroup per (child, group of output rows): sums every live row by column
// class into LDS, then writes the rows with their sums and maxima.  The work
// item's child comes from a binary search of the per-child item prefix.  CB
// threads per workgroup: large children, whose rows go one per workgroup, take
// more threads, so more of a row's loads are in flight per CU.  AccT: the LDS
// accumulator, uint32_t when every node of the level weighs less than 2^32 in
// total (no class sum can exceed it; half the LDS and one bank per atomic),
// unsigned long long otherwise.
template <int CB, typename AccT>
__global__ void __launch_bounds__(CB) k_ks_compact(const KsChild* __restrict__ children,
                                                   const uint32_t* __restrict__ k_out,
                                                   const uint32_t* __restrict__ item_prefix, uint32_t nchild,
                                                   KsPool pool) {
  extern __shared__ __align__(16) unsigned char acc_lds[];
  AccT* acc = reinterpret_cast<AccT*>(acc_lds);
  const uint32_t item = blockIdx.x;
  uint32_t lo = 0, hi = nchild;  // last child with item_prefix[child] <= item
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (item_prefix[mid] <= item) lo = mid;
    else hi = mid;
  }
  const KsChild c = children[lo];
  const uint32_t kk = k_out[lo];
  if (kk == 0 || kk > c.tgt) return;  // disconnected child, or an error code
  const uint32_t R = compact_rows(c.tgt);
  const uint32_t r0 = (item - item_prefix[lo]) * R;
  if (r0 >= kk) return;
  const uint32_t nr = min(R, kk - r0);
  const uint32_t s = c.s;
  for (uint32_t x = threadIdx.x; x < nr * kk; x += CB) acc[x] = 0;
  __syncthreads();
  // Rows and the column map are global memory (the struct's pointers are
  // generic, which would give flat loads); kCompactUnroll elements per thread
  // are loaded before any is added, so a wavefront keeps several 512-byte
  // loads in flight instead of one.
  typedef const __attribute__((address_space(1))) int64_t g_i64;
  typedef const __attribute__((address_space(1))) uint32_t g_u32;
  constexpr uint32_t U = kCompactUnroll;
  g_u32* cm = (g_u32*)c.colmap;
  for (uint32_t rr = 0; rr < nr; ++rr) {
    const uint32_t e = c.live[r0 + rr];
    g_i64* src = (g_i64*)((e & kLiveWork) ? pool.base + (size_t)(e & ~kLiveWork) * kKsUnit : c.parent + (size_t)e * s);
    AccT* a = acc + (size_t)rr * kk;
    uint32_t y = threadIdx.x;
    for (; y + (U - 1) * CB < s; y += U * CB) {
      int64_t v[U];
      uint32_t m[U];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        v[u] = src[y + u * CB];
        m[u] = cm[y + u * CB];
      }
#pragma unroll
      for (uint32_t u = 0; u < U; ++u)
        if (v[u]) atomicAdd(&a[m[u]], (AccT)v[u]);
    }
    for (; y < s; y += CB) {
      const int64_t v = src[y];
      if (v) atomicAdd(&a[cm[y]], (AccT)v);
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t rr = wave; rr < nr; rr += CB / 64) {
    const uint32_t r = r0 + rr;
    int64_t* dst = c.out + (size_t)r * kk;
    int64_t sm = 0, mx = 0;
    for (uint32_t x = lane; x < kk; x += 64) {
      const int64_t v = (int64_t)(uint64_t)acc[(size_t)rr * kk + x];
      dst[x] = v;
      sm += v;
      mx = v > mx ? v : mx;
    }
    // entries are non-negative, so the unsigned maximum is the maximum
    sm = (int64_t)wave_sum<uint64_t>((uint64_t)sm);
    mx = (int64_t)wave_max<uint64_t>((uint64_t)mx);
    if (lane == 0) {
      c.orsum[r] = sm;
      c.ormax[r] = mx;
    }
  }
}

// Row sums of a dense matrix (one wavefront per row) and their minimum.
__global__ void __launch_bounds__(256) k_row_sums(const int64_t* __restrict__ A, uint32_t n, int64_t* __restrict__ rsum,
                                                  int64_t* out_min) {
  __shared__ long long wmin[4];
  long long best = INT64_MAX;
  for (uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6); r < n; r += gridDim.x * 4) {
    const int64_t* row = A + (size_t)r * n;
    uint64_t acc = 0;
    for (uint32_t x = threadIdx.x & 63; x < n; x += 64) acc += (uint64_t)row[x];
    acc = wave_sum<uint64_t>(acc);
    if ((threadIdx.x & 63) == 0) rsum[r] = (int64_t)acc;
    best = (long long)acc < best ? (long long)acc : best;
  }
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long b = wmin[0];
    for (int w = 1; w < 4; ++w) b = wmin[w] < b ? wmin[w] : b;
    atomicMin((long long*)out_min, b);
  }
}

__global__ void k_fill_i64(int64_t* p, uint32_t count, int64_t v) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_zero_words(unsigned long long* p, uint32_t count) {
  for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) p[i] = 0ull;
}

// Block size for a child of s vertices: the smallest power of two from 64 up
// with at most kKsChunk row entries per thread, unless the context holds a
// tuned size for the row-size class of s (cac_tune_set); any power of two from
// that minimum to 1024 gives the same results.
inline int contract_block(const cac_ctx* ctx, uint32_t s) {
  int B = 64;
  while (B < 1024 && (uint32_t)B * kKsChunk < s) B *= 2;
  const int cls = B == 64 ? 0 : B == 128 ? 1 : B == 256 ? 2 : B == 512 ? 3 : -1;
  if (cls >= 0 && ctx->ks_contract_block[cls] >= B) B = ctx->ks_contract_block[cls];
  return B;
}

constexpr uint32_t kFewChildren = 256;  // the CU count of one MI355X

// Bump carving with 64-byte alignment (level buffers hold millions of small arrays).
struct Carve64 {
  char* base;
  size_t off = 0;
  explicit Carve64(void* b) : base((char*)b) {}
  template <class T>
  T* take(size_t count) {
    off = (off + 63) & ~size_t(63);
    T* p = (T*)(base + off);
    off += count * sizeof(T);
    return p;
  }
};
inline size_t c64(size_t bytes) { return (bytes + 63) & ~size_t(63); }

}  // namespace

// Host mirror of co_mincut.cpp:18-20 (same double arithmetic).
uint32_t ks_target_host(uint32_t s) {
  int t = (int)std::ceil((double)s / std::sqrt((double)2)) + 1;
  return t > (int)kKsBase ? (uint32_t)t : kKsBase;
}

namespace {
inline uint32_t ks_target(uint32_t s) { return ks_target_host(s); }

// Grow-only level buffer `which` of the context.
int ks_reserve(cac_ctx* ctx, int which, size_t bytes, void** out) {
  int rc = grow_buffer(ctx, &ctx->ks_buf[which], &ctx->ks_bytes[which], bytes);
  *out = ctx->ks_buf[which];
  return rc;
}

// Device bytes of the per-child arrays of one child (out, sums, maxima, map, list).
size_t child_bytes(uint32_t s, uint32_t t) {
  return c64((size_t)t * t * 8) + 2 * c64((size_t)t * 8) + c64((size_t)s * 4) + c64((size_t)t * 4);
}
size_t child_pool_units(uint32_t s, uint32_t t) { return (size_t)(s - t) * ((s + kKsUnit - 1) / kKsUnit); }

int run_leaves(cac_ctx* ctx, const std::vector<KsLeaf>& leaves, int buf, int64_t* d_results, hipStream_t s) {
  const size_t n = leaves.size();
  void* base = nullptr;
  int rc = ks_reserve(ctx, buf, c64(n * sizeof(KsLeaf)) + c64(n * 4) + 128, &base);
  if (rc) return rc;
  Carve64 cv(base);
  KsLeaf* d_leaves = cv.take<KsLeaf>(n);
  uint32_t* d_list = cv.take<uint32_t>(n);
  uint32_t* d_count = cv.take<uint32_t>(2);  // [0] 32-bit leaves (list front), [1] the others (list back)
  CAC_HIP(ctx, hipMemcpyAsync(d_leaves, leaves.data(), n * sizeof(KsLeaf), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, (unsigned long long*)d_count, 1u);
  prof_begin(ctx, "ks_cert", s);
  hipLaunchKernelGGL(k_ks_cert, dim3((unsigned)n), dim3(64), 0, s, (const KsLeaf*)d_leaves, d_results, d_list, d_count);
  prof_end(ctx, s);
  CAC_LAUNCH_CHECK(ctx, "ks certificate");
  CAC_HIP(ctx, hipMemcpyAsync(ctx->pinned, d_count, 8, hipMemcpyDeviceToHost, s));
  CAC_HIP(ctx, hipStreamSynchronize(s));
  const uint32_t n32 = ((uint32_t*)ctx->pinned)[0], n64 = ((uint32_t*)ctx->pinned)[1];
  if (n32 + n64) {
    prof_begin(ctx, "ks_sw", s);
    if (n32)
      hipLaunchKernelGGL(k_ks_sw<uint32_t>, dim3(n32), dim3(64), 0, s, (const KsLeaf*)d_leaves,
                         (const uint32_t*)d_list, d_results);
    if (n64)
      hipLaunchKernelGGL(k_ks_sw<int64_t>, dim3(n64), dim3(64), 0, s, (const KsLeaf*)d_leaves,
                         (const uint32_t*)(d_list + (n - n64)), d_results);
    prof_end(ctx, s);
    CAC_LAUNCH_CHECK(ctx, "ks stoer-wagner");
  }
  CAC_HIP(ctx, hipStreamSynchronize(s));
  return CAC_OK;
}

template <int B>
void launch_contract(const KsChild* d_children, uint32_t count, uint32_t smax, KsPool pool, uint32_t* d_kout,
                     int64_t* d_results, hipStream_t s) {
  hipLaunchKernelGGL(k_ks_contract<B>, dim3(count), dim3(B), contract_lds_bytes(smax, B), s, d_children, pool, d_kout,
                     d_results);
}

// One level's launches for the children ch (already in d_children's order, by
// block-size class): contraction per class; waits for the children's k_out;
// then the compaction, sized by those k (a child that the PR pass collapsed
// takes a few workgroups, not t' / rows of them), left running on the stream
// (the next level's work follows it there).
int contract_level(cac_ctx* ctx, const std::vector<KsChild>& ch, KsChild* d_children, uint32_t* d_prefix,
                   uint32_t* d_kout, unsigned long long* d_used, KsPool pool, uint32_t max_tgt, int64_t* d_results,
                   std::vector<uint32_t>& kout, bool w32, hipStream_t s) {
  const size_t nchild = ch.size();
  CAC_HIP(ctx, hipMemcpyAsync(d_children, ch.data(), nchild * sizeof(KsChild), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, d_used, 1u);
  prof_begin(ctx, "ks_contract", s);
  size_t c0 = 0;
  while (c0 < nchild) {
    int B = contract_block(ctx, ch[c0].s);
    size_t c1 = c0;
    uint32_t smax = 0;
    while (c1 < nchild && contract_block(ctx, ch[c1].s) == B) smax = std::max(smax, ch[c1++].s);
    const uint32_t cnt = (uint32_t)(c1 - c0);
    // Fewer children than CUs (the top levels of a small solve): the device is
    // idle but for these workgroups, and a wider block shortens each pick's
    // row scans (ER(1000, 0.05) square_root: contraction 1.5 -> 1.33 ms per
    // solve).  With many children the narrow block keeps occupancy up.
    if (cnt < kFewChildren && B < 256) B = 256;
    if (B == 1024) launch_contract<1024>(d_children + c0, cnt, smax, pool, d_kout + c0, d_results, s);
    else if (B == 512) launch_contract<512>(d_children + c0, cnt, smax, pool, d_kout + c0, d_results, s);
    else if (B == 256) launch_contract<256>(d_children + c0, cnt, smax, pool, d_kout + c0, d_results, s);
    else if (B == 128) launch_contract<128>(d_children + c0, cnt, smax, pool, d_kout + c0, d_results, s);
    else launch_contract<64>(d_children + c0, cnt, smax, pool, d_kout + c0, d_results, s);
    c0 = c1;
  }
  prof_end(ctx, s);
  CAC_LAUNCH_CHECK(ctx, "ks contract");
  CAC_HIP(ctx, hipMemcpyAsync(kout.data(), d_kout, nchild * 4, hipMemcpyDeviceToHost, s));
  CAC_HIP(ctx, hipStreamSynchronize(s));
  // compaction work items: ceil(k / rows) per child of 1 <= k <= t' (0 for a
  // disconnected child and for the error codes)
  std::vector<uint32_t>& item_prefix = ctx->ks_items;  // rewritten only after the next synchronisation
  item_prefix.assign(nchild + 1, 0u);
  uint32_t items = 0;
  for (size_t i = 0; i < nchild; ++i) {
    item_prefix[i] = items;
    const uint32_t k = kout[i];
    if (k >= 1 && k <= ch[i].tgt) items += (k + compact_rows(ch[i].tgt) - 1) / compact_rows(ch[i].tgt);
  }
  item_prefix[nchild] = items;
  if (items == 0) return CAC_OK;
  CAC_HIP(ctx, hipMemcpyAsync(d_prefix, item_prefix.data(), (nchild + 1) * 4, hipMemcpyHostToDevice, s));
  prof_begin(ctx, "ks_compact", s);
  const size_t acc_bytes = w32 ? 4 : 8;
  const size_t lds = std::max<size_t>(4096, max_tgt) * acc_bytes;
  const int cb = compact_block(ctx, max_tgt);
  const KsChild* dc = d_children;
  const uint32_t* dk = d_kout;
  const uint32_t* dp = d_prefix;
  const uint32_t nc = (uint32_t)nchild;
  if (w32) {
    if (cb == 1024) hipLaunchKernelGGL((k_ks_compact<1024, uint32_t>), dim3(items), dim3(1024), lds, s, dc, dk, dp, nc, pool);
    else if (cb == 512) hipLaunchKernelGGL((k_ks_compact<512, uint32_t>), dim3(items), dim3(512), lds, s, dc, dk, dp, nc, pool);
    else hipLaunchKernelGGL((k_ks_compact<256, uint32_t>), dim3(items), dim3(256), lds, s, dc, dk, dp, nc, pool);
  } else {
    typedef unsigned long long u64;
    if (cb == 1024) hipLaunchKernelGGL((k_ks_compact<1024, u64>), dim3(items), dim3(1024), lds, s, dc, dk, dp, nc, pool);
    else if (cb == 512) hipLaunchKernelGGL((k_ks_compact<512, u64>), dim3(items), dim3(512), lds, s, dc, dk, dp, nc, pool);
    else hipLaunchKernelGGL((k_ks_compact<256, u64>), dim3(items), dim3(256), lds, s, dc, dk, dp, nc, pool);
  }
  prof_end(ctx, s);
  CAC_LAUNCH_CHECK(ctx, "ks compact");
  return CAC_OK;
}

}  // namespace

int ks_row_sums(cac_ctx* ctx, const int64_t* d_adj, uint32_t n, int64_t* d_rsum, int64_t* d_min, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, s, d_min, 1u, (int64_t)INT64_MAX);
  hipLaunchKernelGGL(k_row_sums, dim3(grid_for(n, 4, 2048)), dim3(256), 0, s, d_adj, n, d_rsum, d_min);
  CAC_LAUNCH_CHECK(ctx, "row sums");
  return CAC_OK;
}

size_t ks_live_bytes_per_root(uint32_t n) {
  // children of every node at one depth with their pool rows (and 1/8 spare);
  // a depth's children go to the buffer its parents are not in, so the two
  // buffers hold the largest even and the largest odd depth
  size_t best[2] = {0, 0};
  uint64_t nodes = 1;
  int depth = 0;
  for (uint32_t sz = n; sz > kKsBase; sz = ks_target(sz), nodes *= 2, ++depth) {
    const uint32_t t = ks_target(sz);
    const size_t b =
        2 * nodes * (child_bytes(sz, t) + sizeof(KsChild) + 16 + child_pool_units(sz, t) * kKsUnit * 8 * 9 / 8);
    best[depth & 1] = std::max(best[depth & 1], b);
  }
  return best[0] + best[1] + 8192;
}

// Runs one Karger-Stein trial per root.  roots[i]: device matrix (n_i x n_i)
// with its row sums, trial id, result slot and bound; d_results[slot] must hold
// the bound already and receives min(bound, every leaf) by atomicMin.
int ks_run(cac_ctx* ctx, const std::vector<KsRoot>& roots, int64_t* d_results, hipStream_t s) {
  if (roots.empty()) return CAC_OK;
  struct Node {
    const int64_t* mat;
    const int64_t* rsum;
    const int64_t* rmax;
    uint32_t s, trial, nid, slot;
    int64_t bound;
    uint32_t seed;
    bool w32;  // the root's total weight is below 2^32 (a child's total never exceeds its parent's)
  };
  static std::once_flag lds_attr;
  hipError_t attr_err = hipSuccess;
  std::call_once(lds_attr, [&] {
    attr_err = hipFuncSetAttribute((const void*)k_ks_contract<1024>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)contract_lds_bytes(kKsMaxS, 1024));
    if (attr_err == hipSuccess)
      attr_err = hipFuncSetAttribute((const void*)k_ks_contract<512>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)contract_lds_bytes(512 * kKsChunk, 512));
  });
  CAC_HIP(ctx, attr_err);
  // parents and children of a level live in the two level buffers
  BufPin pin0(ctx, kBufKs0), pin1(ctx, kBufKs1);

  std::vector<Node> level;
  for (const KsRoot& r : roots) {
    if (r.n > kKsMaxS) return fail(ctx, CAC_EINVAL, "dense Karger-Stein supports at most 8192 vertices");
    level.push_back({r.mat, r.rsum, nullptr, r.n, r.trial, r.nid, r.slot, r.bound, r.seed, r.w32});
  }
  int buf = 0;  // level buffer for the children of this level; the parents are in the other one
  while (!level.empty()) {
    std::vector<KsLeaf> leaves;
    std::vector<Node> internal;
    for (auto& nd : level) {
      if (nd.s <= kKsBase) leaves.push_back({nd.mat, nd.rsum, nd.rmax, nd.s, nd.slot, nd.bound});
      else internal.push_back(nd);
    }
    if (!leaves.empty()) {
      int rc = run_leaves(ctx, leaves, buf, d_results, s);
      if (rc) return rc;
    }
    if (internal.empty()) break;
    // children sorted by block-size class (1024, 256, 64), two per internal node
    std::vector<std::pair<int, uint32_t>> order;
    for (uint32_t i = 0; i < internal.size(); ++i) order.push_back({-contract_block(ctx, internal[i].s), i});
    std::stable_sort(order.begin(), order.end());
    const size_t nchild = 2 * internal.size();
    size_t arrays = 0, units = 0, max_units_child = 0;
    uint32_t max_tgt = 0;
    for (auto& nd : internal) {
      const uint32_t t = ks_target(nd.s);
      arrays += 2 * child_bytes(nd.s, t);
      units += 2 * child_pool_units(nd.s, t);
      max_units_child = std::max<size_t>(max_units_child, (nd.s + kKsUnit - 1) / kKsUnit * (size_t)nd.s);
      max_tgt = std::max(max_tgt, t);
    }
    std::vector<KsChild> ch(nchild);
    std::vector<uint32_t> item_prefix(nchild + 1);
    std::vector<uint32_t> kout(nchild);
    // first pass: pool rows for the picks with 1/8 spare; children whose PR pass
    // needs more come back as kKsOverflow and are redone below
    const size_t pool_units = units + std::max<size_t>(units / 8, max_units_child);
    if (pool_units >= kLiveWork) return fail(ctx, CAC_ENOMEM, "Karger-Stein level too large for one pass");
    const size_t bytes = c64(nchild * sizeof(KsChild)) + c64((nchild + 1) * 4) + c64(nchild * 4) + 64 + arrays +
                         pool_units * kKsUnit * 8 + 1024;
    void* base = nullptr;
    int rc = ks_reserve(ctx, buf, bytes, &base);
    if (rc) return rc;
    Carve64 cv(base);
    KsChild* d_children = cv.take<KsChild>(nchild);
    uint32_t* d_prefix = cv.take<uint32_t>(nchild + 1);
    uint32_t* d_kout = cv.take<uint32_t>(nchild);
    unsigned long long* d_used = cv.take<unsigned long long>(1);
    size_t ci = 0;
    for (auto& o : order) {
      const Node& nd = internal[o.second];
      const uint32_t t = ks_target(nd.s);
      for (uint32_t k = 0; k < 2; ++k, ++ci) {
        KsChild& x = ch[ci];
        x.parent = nd.mat;
        x.prsum = nd.rsum;
        x.out = cv.take<int64_t>((size_t)t * t);
        x.orsum = cv.take<int64_t>(t);
        x.ormax = cv.take<int64_t>(t);
        x.colmap = cv.take<uint32_t>(nd.s);
        x.live = cv.take<uint32_t>(t);
        x.s = nd.s;
        x.tgt = t;
        x.trial = nd.trial;
        x.nid = 2 * nd.nid + k;
        x.slot = nd.slot;
        x.seed = nd.seed;
        x.bound = nd.bound;
      }
    }
    cv.off = (cv.off + 511) & ~size_t(511);
    KsPool pool{(int64_t*)(cv.base + cv.off), d_used, pool_units};
    bool w32 = true;
    for (const Node& nd : internal) w32 = w32 && nd.w32;
    rc = contract_level(ctx, ch, d_children, d_prefix, d_kout, d_used, pool, max_tgt, d_results, kout, w32, s);
    if (rc) return rc;
    // children that ran out of pool rows: again, with s pool rows each (a child
    // merges at most s - 2 times, so this cannot overflow)
    std::vector<uint32_t> over;
    for (uint32_t i = 0; i < nchild; ++i)
      if (kout[i] == kKsOverflow) over.push_back(i);
    if (!over.empty()) {
      std::vector<KsChild> rch;
      size_t runits = 0;
      for (uint32_t i : over) {
        rch.push_back(ch[i]);
        runits += (size_t)ch[i].s * ((ch[i].s + kKsUnit - 1) / kKsUnit);
      }
      const size_t no = rch.size();
      if (runits >= kLiveWork) return fail(ctx, CAC_ENOMEM, "Karger-Stein level too large for one pass");
      const size_t rbytes = c64(no * sizeof(KsChild)) + c64((no + 1) * 4) + c64(no * 4) + 64 + runits * kKsUnit * 8 + 1024;
      rc = grow_buffer(ctx, &ctx->ks_retry_buf, &ctx->ks_retry_bytes, rbytes);
      if (rc) return rc;
      Carve64 rv(ctx->ks_retry_buf);
      KsChild* d_rch = rv.take<KsChild>(no);
      uint32_t* d_rprefix = rv.take<uint32_t>(no + 1);
      uint32_t* d_rkout = rv.take<uint32_t>(no);
      unsigned long long* d_rused = rv.take<unsigned long long>(1);
      rv.off = (rv.off + 511) & ~size_t(511);
      KsPool rpool{(int64_t*)(rv.base + rv.off), d_rused, runits};
      std::vector<uint32_t> rkout(no);
      rc = contract_level(ctx, rch, d_rch, d_rprefix, d_rkout, d_rused, rpool, max_tgt, d_results, rkout, w32, s);
      if (rc) return rc;
      for (size_t k = 0; k < no; ++k) {
        if (rkout[k] == kKsOverflow) return fail(ctx, CAC_EDEVICE, "Karger-Stein row pool overflow");
        kout[over[k]] = rkout[k];
      }
    }
    std::vector<Node> next;
    for (size_t i = 0; i < nchild; ++i) {
      if (kout[i] == kKsBad) return fail(ctx, CAC_EDEVICE, "Karger-Stein selection out of range");
      if (kout[i] == 0) continue;  // disconnected child: its 0 is already recorded
      next.push_back({ch[i].out, ch[i].orsum, ch[i].ormax, kout[i], ch[i].trial, ch[i].nid, ch[i].slot, ch[i].bound,
                      ch[i].seed, w32});
    }
    level.swap(next);
    buf ^= 1;
  }
  CAC_HIP(ctx, hipStreamSynchronize(s));
  return CAC_OK;
}

// Validates a host matrix: symmetric, zero diagonal, non-negative, no overflow.
int ks_check_matrix(cac_ctx* ctx, const int64_t* adj, uint32_t n) {
  long double total = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (adj[(size_t)i * n + i] != 0) return fail(ctx, CAC_EINVAL, "adjacency diagonal must be zero");
    for (uint32_t j = 0; j < n; ++j) {
      const int64_t v = adj[(size_t)i * n + j];
      if (v < 0) return fail(ctx, CAC_EINVAL, "negative edge weight");
      if (j > i && v != adj[(size_t)j * n + i]) return fail(ctx, CAC_EINVAL, "adjacency matrix is not symmetric");
      total += v;
    }
  }
  if (total >= 4.0e18L) return fail(ctx, CAC_ERANGE, "total weight exceeds int64 range");
  return CAC_OK;
}

int ks_dense_run(cac_ctx* ctx, const int64_t* d_adj, uint32_t n, uint32_t seed, uint32_t first_trial,
                 uint32_t ntrials, int64_t* raw_cuts, int64_t* min_cut, hipStream_t s) {
  if (ntrials == 0) return fail(ctx, CAC_EINVAL, "ntrials must be >= 1");
  if (n > kKsMaxS) return fail(ctx, CAC_EINVAL, "dense Karger-Stein supports at most 8192 vertices");
  int64_t* d_small = nullptr;  // [0] bound, [1..n] row sums, then the trial results
  CAC_HIP(ctx, hipMalloc((void**)&d_small, 8 * ((size_t)n + ntrials + 2)));
  int64_t* d_bound = d_small;
  int64_t* d_rsum = d_small + 1;
  int64_t* d_res = d_small + 1 + n;
  int rc = ks_row_sums(ctx, d_adj, n, d_rsum, d_bound, s);
  int64_t bound = 0;
  std::vector<int64_t> rs(n);  // the row sums: their total decides 32-bit compaction sums
  if (rc == CAC_OK) {
    hipError_t e = hipMemcpyAsync(&bound, d_bound, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(rs.data(), d_rsum, 8 * (size_t)n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(ctx, e, "row sums");
  }
  bool w32 = rc == CAC_OK;  // total weight below 2^32, checked without wrapping
  uint64_t total = 0;
  for (int64_t r : rs) {
    if (r < 0 || (uint64_t)r >= (1ull << 32) || (total += (uint64_t)r) >= (1ull << 32)) {
      w32 = false;
      break;
    }
  }
  if (rc == CAC_OK) {
    hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(ntrials, 256, 64)), dim3(256), 0, s, d_res, ntrials, bound);
    // trials run in batches whose two live levels fit in 3/4 of the free
    // memory: the top levels are chains of picks with one workgroup per child,
    // so the more trials a batch holds, the more CUs they keep busy.  Every
    // batch but the last is as large as that allows: at n = 8192 the first of
    // two then fills the CUs at the levels where children outnumber them (two
    // even batches of 85 left a partial round of workgroups in each)
    size_t avail = 0;
    rc = device_memory_available(ctx, &avail);
    const size_t per_trial = ks_live_bytes_per_root(n);
    size_t batch = std::max<size_t>(1, (size_t)(0.75 * (double)avail) / per_trial);
    batch = std::min<size_t>(batch, ntrials);
    for (uint32_t t0 = 0; t0 < ntrials && rc == CAC_OK; t0 += (uint32_t)batch) {
      std::vector<KsRoot> roots;
      for (uint32_t t = t0; t < std::min<uint32_t>(ntrials, t0 + (uint32_t)batch); ++t)
        roots.push_back({d_adj, d_rsum, n, first_trial + t, t, bound, seed, 1, w32});
      rc = ks_run(ctx, roots, d_res, s);
    }
  }
  if (rc == CAC_OK) {
    std::vector<int64_t> res(ntrials);
    hipError_t e = hipMemcpyAsync(res.data(), d_res, 8 * (size_t)ntrials, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(ctx, e, "trial results");
    int64_t mn = INT64_MAX;
    for (uint32_t t = 0; t < ntrials && rc == CAC_OK; ++t) {
      if (raw_cuts) raw_cuts[t] = res[t];
      mn = std::min(mn, res[t]);
    }
    if (rc == CAC_OK && min_cut) *min_cut = mn;
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(d_small);
  return rc;
}

}  // namespace cac

extern "C" {

int cac_ks_dense_dev(cac_ctx* ctx, const int64_t* d_adj, uint32_t n, uint32_t seed, uint32_t first_trial,
                     uint32_t ntrials, int64_t* raw_cuts, int64_t* min_cut, void* stream) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if (!d_adj || n < 2) return cac::fail(ctx, CAC_EINVAL, "need a matrix with at least 2 vertices");
  return cac::ks_dense_run(ctx, d_adj, n, seed, first_trial, ntrials, raw_cuts, min_cut, cac::pick_stream(ctx, stream));
}

int cac_ks_dense(cac_ctx* ctx, const int64_t* adj, uint32_t n, uint32_t seed, uint32_t first_trial, uint32_t ntrials,
                 int64_t* raw_cuts, int64_t* min_cut) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if (!adj || n < 2) return cac::fail(ctx, CAC_EINVAL, "need a matrix with at least 2 vertices");
  rc = cac::ks_check_matrix(ctx, adj, n);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  int64_t* d_adj = nullptr;
  CAC_HIP(ctx, hipMalloc((void**)&d_adj, (size_t)n * n * 8));
  hipError_t e = hipMemcpyAsync(d_adj, adj, (size_t)n * n * 8, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) {
    (void)hipFree(d_adj);
    return cac::hip_fail(ctx, e, "upload matrix");
  }
  rc = cac::ks_dense_run(ctx, d_adj, n, seed, first_trial, ntrials, raw_cuts, min_cut, s);
  (void)hipStreamSynchronize(s);
  (void)hipFree(d_adj);
  return rc;
}

}  // extern "C"
