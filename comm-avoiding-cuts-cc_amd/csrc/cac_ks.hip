// cac_ks.hip -- dense Karger-Stein recursive contraction on gfx950.
//
// Replaces comincut::minimum_cut / minimum_cut_try on dense graphs
// (src/karger-stein/co_mincut.h:315-390, 466-490) and the recursive-contract base
// case mincut::parallel_cut hands to it (src/recursive-contract/recursive_contract.cpp:262-266).
//
// Execution model: the recursion tree of every trial is walked breadth first.  A
// level launch contracts every child of every live node at once (one workgroup
// per child, so trials x 2^level workgroups); the host then gathers the leaves
// (<= 128 vertices) and launches one wavefront per leaf.  PRNG draws are keyed
// by (seed, trial, node id, pick), so the breadth-first order gives bit-identical
// results to the depth-first CPU restatement.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "cac_internal.h"
#include "cac_ks.h"
#include "cac_ks_kernels.h"

namespace cac {
namespace {

// Byte offset of the 2 x kLeafMax int64 leaf-certificate scratch in dynamic LDS.
__host__ __device__ inline size_t leaf_scratch_offset(uint32_t s) {
  return ((13 * (size_t)s + 15) & ~(size_t)15) + 16 * kKsWaves + 64;
}

struct KsLds {
  int64_t* cap;      // [s]  capacities; rank/colmap scratch during compaction
  uint32_t* rep;     // [s]  representatives; colmap during compaction
  uint8_t* flags;    // [s]  bit0: row lives in work; bit1: live at compaction
  uint64_t* red;     // [kKsWaves] reduction scratch
  uint32_t* idx;     // [kKsWaves] index scratch
};

// ---- leaves --------------------------------------------------------------------
// One wavefront per leaf: lane l owns vertices l and l + 64.
constexpr int kLeafMax = 128;

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

__device__ inline int64_t wave_sum_i64(int64_t x) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

// d - (q-1) M, saturating to a large negative value instead of overflowing
// (a negative term only weakens the bound).
__device__ inline int64_t leaf_f(int64_t d, int64_t M, uint32_t q) {
  const int64_t k = (int64_t)q - 1;
  if (M != 0 && k > INT64_MAX / 256 / M) return INT64_MIN / 256;
  return d - k * M;
}

// Lower bound on the minimum cut (see leaf kernel): for |S| = q <= V/2,
// cut(S) >= sum of the q smallest f_q(u) = d(u) - (q-1) M(u).
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
          // keep min when (lower == up)
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

__device__ inline const int64_t* row_of(const KsChild& c, const KsLds& L, uint32_t r) {
  return (L.flags[r] & 1u) ? c.work + (size_t)r * c.s : c.parent + (size_t)r * c.s;
}

// Merge row a into row b (a dies).  Returns the contracted weight
// (cap[a] + cap[b] - cap_new) / 2.  Called by the whole block.
__device__ int64_t ks_merge(const KsChild& c, const KsLds& L, uint32_t a, uint32_t b) {
  const uint32_t s = c.s;
  // every wave has read the representatives that chose (a, b) before any is relabelled
  __syncthreads();
  for (uint32_t x = threadIdx.x; x < s; x += kKsBlock)
    if (L.rep[x] == a) L.rep[x] = b;
  __syncthreads();
  const int64_t* ra = row_of(c, L, a);
  const int64_t* rb = row_of(c, L, b);
  int64_t* dst = c.work + (size_t)b * s;
  uint64_t sum = 0;
#pragma unroll 8
  for (uint32_t x = threadIdx.x; x < s; x += kKsBlock) {
    int64_t v = rb[x] + ra[x];
    if (L.rep[x] == b) v = 0;
    dst[x] = v;
    sum += (uint64_t)v;
  }
  // the new row b is read by other waves in later picks and the compaction
  block_sync_global();
  const int64_t newcap = (int64_t)block_sum_u64(sum, L.red);
  const int64_t w = (L.cap[a] + L.cap[b] - newcap) / 2;
  __syncthreads();
  if (threadIdx.x == 0) {
    L.cap[b] = newcap;
    L.cap[a] = 0;
    L.flags[b] |= 1u;
  }
  __syncthreads();
  return w;
}

__global__ void __launch_bounds__(kKsBlock) k_ks_contract(const KsChild* __restrict__ children, uint32_t seed,
                                                          uint32_t* __restrict__ k_out, int64_t* __restrict__ results) {
  extern __shared__ __align__(16) unsigned char smem[];
  const KsChild c = children[blockIdx.x];
  const uint32_t s = c.s;
  KsLds L;
  L.cap = (int64_t*)smem;
  L.rep = (uint32_t*)(smem + 8 * (size_t)s);
  L.flags = smem + 12 * (size_t)s;
  L.red = (uint64_t*)(smem + ((13 * (size_t)s + 15) & ~(size_t)15));
  L.idx = (uint32_t*)(L.red + kKsWaves);

  // capacities = row sums of the parent; representatives = identity
  for (uint32_t x = threadIdx.x; x < s; x += kKsBlock) {
    L.rep[x] = x;
    L.flags[x] = 0;
  }
  uint64_t W = 0;
  for (uint32_t r = threadIdx.x >> 6; r < s; r += kKsWaves) {
    const int64_t* row = c.parent + (size_t)r * s;
    uint64_t acc = 0;
    for (uint32_t x = threadIdx.x & 63; x < s; x += 64) acc += (uint64_t)row[x];
    acc = wave_sum_u64(acc);
    if ((threadIdx.x & 63) == 0) L.cap[r] = (int64_t)acc;
  }
  __syncthreads();
  {
    uint64_t part = 0;
    for (uint32_t x = threadIdx.x; x < s; x += kKsBlock) part += (uint64_t)L.cap[x];
    W = block_sum_u64(part, L.red);
  }

  uint32_t nv = s, pr1 = 0;
  bool disconnected = false, bad = false;
  for (uint32_t p = 0; p + c.tgt < s; ++p) {
    if (W == 0) {
      disconnected = true;
      break;
    }
    const u32x4 rnd = philox4x32(c.trial, c.nid, p, 0u, seed, kDomKs);
    const uint32_t i = block_select([&](uint32_t x) { return L.cap[x]; }, s, uniform1(lo64(rnd), W), L.red, L.idx);
    if (i >= s) {  // capacities inconsistent with W: never expected; abandon this child
      bad = true;
      break;
    }
    const int64_t* rowi = row_of(c, L, i);
    const uint64_t ci = (uint64_t)L.cap[i];
    const uint32_t k = block_select_row(rowi, s, uniform1(hi64(rnd), ci), L.red, L.idx);
    if (k >= s) {
      bad = true;
      break;
    }
    const uint32_t j = L.rep[k];
    const uint32_t a = i > j ? i : j, b = i > j ? j : i;
    const int64_t w = ks_merge(c, L, a, b);
    W -= 2 * (uint64_t)w;
    --nv;
    if (w >= c.bound) ++pr1;
  }
  if (disconnected || bad) {
    if (threadIdx.x == 0) {
      k_out[blockIdx.x] = bad ? 0xFFFFFFFFu : 0u;
      if (disconnected) atomicMin((long long*)&results[c.slot], 0ll);
    }
    return;
  }

  // PR pass (indexed_adjacency_matrix::pr_pass): contract edges of weight >= bound
  if (pr1 > 0 && c.bound > 0) {
    for (uint32_t i = 0; i < s && nv > 2; ++i) {
      if (L.rep[i] != i || L.cap[i] < c.bound) continue;
      const int64_t* row = row_of(c, L, i);
      uint32_t first = 0xFFFFFFFFu;
      for (uint32_t x = threadIdx.x; x < s; x += kKsBlock)
        if (row[x] >= c.bound) {
          first = x;
          break;
        }
      first = block_min_u32(first, L.idx);
      if (first != 0xFFFFFFFFu) {
        const uint32_t b = L.rep[first];
        (void)ks_merge(c, L, i, b);
        --nv;
      }
    }
  }

  // ---- compaction: rank live rows, sum columns by class --------------------
  __syncthreads();  // the PR loop's reads of cap/rep precede the overwrites below
  uint32_t* rank = (uint32_t*)L.cap;           // [s] in the first half of cap
  uint32_t* colmap_tmp = rank + s;             // [s] in the second half of cap
  {
    // block exclusive scan of live flags, 4-wave chunks of kKsBlock
    uint32_t carry = 0;
    for (uint32_t base = 0; base < s; base += kKsBlock) {
      const uint32_t x = base + threadIdx.x;
      const uint32_t live = (x < s && L.rep[x] == x) ? 1u : 0u;
      const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
      uint32_t inc = live;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
      }
      if (lane == 63) L.idx[wave] = inc;
      __syncthreads();
      uint32_t off = carry, tot = 0;
      for (int w = 0; w < kKsWaves; ++w) {
        uint32_t v = L.idx[w];
        if (w < wave) off += v;
        tot += v;
      }
      if (x < s) {
        rank[x] = off + inc - live;
        if (live) L.flags[x] |= 2u;
      }
      carry += tot;
      __syncthreads();
    }
    nv = carry;
  }
  for (uint32_t y = threadIdx.x; y < s; y += kKsBlock) colmap_tmp[y] = rank[L.rep[y]];
  __syncthreads();
  uint32_t* colmap = L.rep;
  for (uint32_t y = threadIdx.x; y < s; y += kKsBlock) colmap[y] = colmap_tmp[y];
  __syncthreads();
  unsigned long long* acc = (unsigned long long*)L.cap;  // [nv] accumulators
  const uint32_t kk = nv;
  uint32_t out_row = 0;
  for (uint32_t a = 0; a < s; ++a) {
    if (!(L.flags[a] & 2u)) continue;
    for (uint32_t x = threadIdx.x; x < kk; x += kKsBlock) acc[x] = 0ull;
    __syncthreads();
    const int64_t* row = row_of(c, L, a);
    for (uint32_t y = threadIdx.x; y < s; y += kKsBlock) {
      const int64_t v = row[y];
      if (v) atomicAdd(&acc[colmap[y]], (unsigned long long)v);
    }
    __syncthreads();
    int64_t* dst = c.out + (size_t)out_row * kk;
    for (uint32_t x = threadIdx.x; x < kk; x += kKsBlock) dst[x] = (int64_t)acc[x];
    __syncthreads();
    ++out_row;
  }
  // A child that is already a leaf is checked right here: row sums and maxima of
  // its kk x kk output feed the lower-bound certificate; when it proves every cut
  // of the leaf is >= bound the leaf contributes exactly the bound and needs no
  // Stoer-Wagner pass (flag bit 31 of k_out).
  uint32_t flag = 0;
  if (kk <= kLeafMax && kk >= 2) {
    block_sync_global();
    int64_t* dl = (int64_t*)(smem + leaf_scratch_offset(s));
    int64_t* ml = dl + kLeafMax;
    for (uint32_t r = threadIdx.x >> 6; r < kk; r += kKsWaves) {
      const int64_t* row = c.out + (size_t)r * kk;
      int64_t sm = 0, mx = 0;
      for (uint32_t x = threadIdx.x & 63; x < kk; x += 64) {
        const int64_t v = row[x];
        sm += v;
        mx = v > mx ? v : mx;
      }
#pragma unroll
      for (int d = 32; d > 0; d >>= 1) {
        sm += __shfl_xor(sm, d, 64);
        const int64_t o = __shfl_xor(mx, d, 64);
        mx = o > mx ? o : mx;
      }
      if ((threadIdx.x & 63) == 0) {
        dl[r] = sm;
        ml[r] = mx;
      }
    }
    __syncthreads();
    if (threadIdx.x < 64) {
      const int lane = threadIdx.x;
      const bool alive[2] = {(uint32_t)lane < kk, (uint32_t)lane + 64 < kk};
      const int64_t d2[2] = {alive[0] ? dl[lane] : 0, alive[1] ? dl[lane + 64] : 0};
      const int64_t m2[2] = {alive[0] ? ml[lane] : 0, alive[1] ? ml[lane + 64] : 0};
      const int64_t lb = leaf_lower_bound(d2, m2, alive, kk, c.bound);
      if (lb >= c.bound) flag = 0x80000000u;
    }
  }
  if (threadIdx.x == 0) k_out[blockIdx.x] = kk | flag;
}

__global__ void __launch_bounds__(64) k_ks_leaf(const KsLeaf* __restrict__ leaves, int64_t* __restrict__ results) {
  __shared__ int64_t A[kLeafMax * kLeafMax];
  const KsLeaf lf = leaves[blockIdx.x];
  const uint32_t V = lf.k;
  const int lane = threadIdx.x;
  // load (row stride kLeafMax), degrees and row maxima
  int64_t d[2] = {0, 0}, M[2] = {0, 0};
  for (uint32_t r = 0; r < V; ++r) {
    for (uint32_t x = lane; x < V; x += 64) A[r * kLeafMax + x] = lf.mat[(size_t)r * V + x];
  }
  __syncthreads();
  bool alive[2] = {(uint32_t)lane < V, (uint32_t)(lane + 64) < V};
  for (int h = 0; h < 2; ++h) {
    const uint32_t u = lane + 64 * h;
    if (!alive[h]) continue;
    int64_t s = 0, mx = 0;
    for (uint32_t x = 0; x < V; ++x) {
      int64_t v = A[u * kLeafMax + x];
      s += v;
      mx = v > mx ? v : mx;
    }
    d[h] = s;
    M[h] = mx;
  }
  int64_t best;
  const int64_t lb = V >= 2 ? leaf_lower_bound(d, M, alive, V, lf.bound) : 0;
  if (lb >= lf.bound) {
    best = lf.bound;  // every cut of the leaf is at least the bound
  } else {
    // Stoer-Wagner: maximum adjacency order, cut of the phase, merge last two
    best = INT64_MAX;
    uint32_t nalive = V;
    while (nalive > 1) {
      int64_t key[2] = {0, 0};
      bool added[2] = {!alive[0], !alive[1]};
      // start from the smallest live vertex
      uint64_t live_lo = __ballot(alive[0]), live_hi = __ballot(alive[1]);
      uint32_t start = live_lo ? (uint32_t)(__ffsll((long long)live_lo) - 1)
                               : 64u + (uint32_t)(__ffsll((long long)live_hi) - 1);
      uint32_t prev = start, last = start;
      if ((uint32_t)lane == start) added[0] = true;
      if ((uint32_t)lane + 64 == start) added[1] = true;
      for (int h = 0; h < 2; ++h) key[h] += A[start * kLeafMax + lane + 64 * h];
      int64_t cut = 0;
      for (uint32_t step = 1; step < nalive; ++step) {
        int64_t bk = INT64_MIN;
        uint32_t bi = 0xFFFFFFFFu;
        for (int h = 0; h < 2; ++h) {
          if (!added[h] && (key[h] > bk || (key[h] == bk && (uint32_t)lane + 64 * h < bi))) {
            bk = key[h];
            bi = lane + 64 * h;
          }
        }
        wave_argmax(bk, bi);
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
          if (!added[h]) key[h] += A[bi * kLeafMax + lane + 64 * h];
      }
      best = cut < best ? cut : best;
      // merge 'last' into 'prev' (row and column)
      for (int h = 0; h < 2; ++h) {
        const uint32_t x = lane + 64 * h;
        if (x < V && x != prev && x != last) {
          const int64_t v = A[prev * kLeafMax + x] + A[last * kLeafMax + x];
          A[prev * kLeafMax + x] = v;
          A[x * kLeafMax + prev] = v;
        }
      }
      __syncthreads();
      if (lane == 0) {
        A[prev * kLeafMax + last] = 0;
        A[last * kLeafMax + prev] = 0;
      }
      if ((uint32_t)lane == last) alive[0] = false;
      if ((uint32_t)lane + 64 == last) alive[1] = false;
      __syncthreads();
      --nalive;
    }
    best = best < lf.bound ? best : lf.bound;
  }
  if (lane == 0) atomicMin((long long*)&results[lf.slot], (long long)best);
}

__global__ void __launch_bounds__(256) k_min_degree(const int64_t* __restrict__ A, uint32_t n, int64_t* out_bound) {
  // one wave per row, block-level min, one atomic per block
  __shared__ long long wmin[4];
  long long best = INT64_MAX;
  for (uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6); r < n; r += gridDim.x * 4) {
    const int64_t* row = A + (size_t)r * n;
    uint64_t acc = 0;
    for (uint32_t x = threadIdx.x & 63; x < n; x += 64) acc += (uint64_t)row[x];
    acc = wave_sum_u64(acc);
    best = (long long)acc < best ? (long long)acc : best;
  }
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long b = wmin[0];
    for (int w = 1; w < 4; ++w) b = wmin[w] < b ? wmin[w] : b;
    atomicMin((long long*)out_bound, b);
  }
}

__global__ void k_fill_i64(int64_t* p, uint32_t count, int64_t v) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace

// Host mirror of co_mincut.cpp:18-20 (same double arithmetic).
uint32_t ks_target_host(uint32_t s) {
  int t = (int)std::ceil((double)s / std::sqrt((double)2)) + 1;
  return t > (int)kKsBase ? (uint32_t)t : kKsBase;
}

namespace {
size_t contract_lds_bytes(uint32_t s) { return leaf_scratch_offset(s) + 2 * kLeafMax * 8; }
inline uint32_t ks_target(uint32_t s) { return ks_target_host(s); }
}  // namespace

int ks_min_degree(cac_ctx* ctx, const int64_t* d_adj, uint32_t n, int64_t* d_bound, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, s, d_bound, 1u, (int64_t)INT64_MAX);
  hipLaunchKernelGGL(k_min_degree, dim3(grid_for(n, 4, 2048)), dim3(256), 0, s, d_adj, n, d_bound);
  CAC_LAUNCH_CHECK(ctx, "min degree");
  return CAC_OK;
}

// Runs one Karger-Stein trial per root.  roots[i]: device matrix (n_i x n_i),
// trial id, result slot and bound; d_results[slot] must hold the bound already.
int ks_run(cac_ctx* ctx, const std::vector<KsRoot>& roots, uint32_t seed, int64_t* d_results, hipStream_t s) {
  if (roots.empty()) return CAC_OK;
  struct Node {
    const int64_t* mat;
    uint32_t s, trial, nid, slot;
    int64_t bound;
  };
  std::vector<Node> level;
  for (size_t i = 0; i < roots.size(); ++i)
    level.push_back({roots[i].mat, roots[i].n, roots[i].trial, 1u, roots[i].slot, roots[i].bound});

  void* prev_arena = nullptr;  // buffers holding 'level' matrices (except the roots)
  int rc = CAC_OK;
  while (!level.empty() && rc == CAC_OK) {
    std::vector<KsLeaf> leaves;
    std::vector<Node> internal;
    for (auto& nd : level) {
      if (nd.s <= kKsBase) leaves.push_back({nd.mat, nd.s, nd.slot, nd.bound});
      else internal.push_back(nd);
    }
    if (!leaves.empty()) {
      KsLeaf* d_leaves = nullptr;
      hipError_t e = hipMalloc((void**)&d_leaves, leaves.size() * sizeof(KsLeaf));
      if (e != hipSuccess) { rc = hip_fail(ctx, e, "leaf descriptors"); break; }
      CAC_HIP(ctx, hipMemcpyAsync(d_leaves, leaves.data(), leaves.size() * sizeof(KsLeaf), hipMemcpyHostToDevice, s));
      prof_begin(ctx, "ks_leaf", s);
      hipLaunchKernelGGL(k_ks_leaf, dim3((unsigned)leaves.size()), dim3(64), 0, s, d_leaves, d_results);
      prof_end(ctx, s);
      CAC_LAUNCH_CHECK(ctx, "ks leaf");
      CAC_HIP(ctx, hipStreamSynchronize(s));
      CAC_HIP(ctx, hipFree(d_leaves));
    }
    if (internal.empty()) break;
    // children: 2 per internal node
    size_t bytes = 0;
    uint32_t smax = 0;
    for (auto& nd : internal) {
      const uint32_t t = ks_target(nd.s);
      bytes += 2 * (carve_size((size_t)nd.s * nd.s * 8) + carve_size((size_t)t * t * 8));
      smax = std::max(smax, nd.s);
    }
    if (smax > kKsMaxLdsS) { rc = fail(ctx, CAC_EINVAL, "dense Karger-Stein supports at most 8192 vertices"); break; }
    const size_t nchild = 2 * internal.size();
    bytes += carve_size(nchild * sizeof(KsChild)) + carve_size(nchild * 4);
    void* arena = nullptr;
    hipError_t e = hipMalloc(&arena, bytes);
    if (e != hipSuccess) { rc = hip_fail(ctx, e, "Karger-Stein level buffers"); break; }
    Carver cv(arena);
    KsChild* d_children = cv.take<KsChild>(nchild);
    uint32_t* d_kout = cv.take<uint32_t>(nchild);
    std::vector<KsChild> ch(nchild);
    for (size_t i = 0; i < internal.size(); ++i) {
      const Node& nd = internal[i];
      const uint32_t t = ks_target(nd.s);
      for (uint32_t c = 0; c < 2; ++c) {
        KsChild& x = ch[2 * i + c];
        x.parent = nd.mat;
        x.work = cv.take<int64_t>((size_t)nd.s * nd.s);
        x.out = cv.take<int64_t>((size_t)t * t);
        x.s = nd.s;
        x.tgt = t;
        x.trial = nd.trial;
        x.nid = 2 * nd.nid + c;
        x.slot = nd.slot;
        x.pad = 0;
        x.bound = nd.bound;
      }
    }
    CAC_HIP(ctx, hipMemcpyAsync(d_children, ch.data(), nchild * sizeof(KsChild), hipMemcpyHostToDevice, s));
    static std::once_flag lds_attr;
    hipError_t attr_err = hipSuccess;
    std::call_once(lds_attr, [&] {
      attr_err = hipFuncSetAttribute((const void*)k_ks_contract, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)contract_lds_bytes(kKsMaxLdsS));
    });
    CAC_HIP(ctx, attr_err);
    prof_begin(ctx, "ks_contract", s);
    hipLaunchKernelGGL(k_ks_contract, dim3((unsigned)nchild), dim3(kKsBlock), contract_lds_bytes(smax), s, d_children,
                       seed, d_kout, d_results);
    prof_end(ctx, s);
    CAC_LAUNCH_CHECK(ctx, "ks contract");
    std::vector<uint32_t> kout(nchild);
    CAC_HIP(ctx, hipMemcpyAsync(kout.data(), d_kout, nchild * 4, hipMemcpyDeviceToHost, s));
    CAC_HIP(ctx, hipStreamSynchronize(s));
    if (prev_arena) CAC_HIP(ctx, hipFree(prev_arena));
    prev_arena = arena;
    std::vector<Node> next;
    for (size_t i = 0; i < nchild; ++i) {
      if (kout[i] == 0xFFFFFFFFu) rc = fail(ctx, CAC_EDEVICE, "Karger-Stein selection out of range");
      if (kout[i] == 0 || kout[i] == 0xFFFFFFFFu) continue;  // disconnected child: its 0 is already recorded
      if (kout[i] & 0x80000000u) continue;                     // certified leaf: contributes the bound
      next.push_back({ch[i].out, kout[i], ch[i].trial, ch[i].nid, ch[i].slot, ch[i].bound});
    }
    level.swap(next);
  }
  CAC_HIP(ctx, hipStreamSynchronize(s));
  if (prev_arena) CAC_HIP(ctx, hipFree(prev_arena));
  return rc;
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
  int64_t *d_bound = nullptr, *d_res = nullptr;
  CAC_HIP(ctx, hipMalloc((void**)&d_bound, 8 * (size_t)(ntrials + 1)));
  CAC_HIP(ctx, hipMalloc((void**)&d_res, 8 * (size_t)ntrials));
  int rc = ks_min_degree(ctx, d_adj, n, d_bound, s);
  if (rc) return rc;
  int64_t bound = 0;
  CAC_HIP(ctx, hipMemcpyAsync(&bound, d_bound, 8, hipMemcpyDeviceToHost, s));
  CAC_HIP(ctx, hipStreamSynchronize(s));
  std::vector<int64_t> hb(ntrials, bound);
  CAC_HIP(ctx, hipMemcpyAsync(d_bound, hb.data(), 8 * (size_t)ntrials, hipMemcpyHostToDevice, s));
  CAC_HIP(ctx, hipMemcpyAsync(d_res, hb.data(), 8 * (size_t)ntrials, hipMemcpyHostToDevice, s));
  // trials run in batches sized to the free device memory
  size_t free_b = 0, total_b = 0;
  CAC_HIP(ctx, hipMemGetInfo(&free_b, &total_b));
  size_t per_trial = 0;
  {
    uint64_t nodes = 1;
    for (uint32_t sz = n; sz > kKsBase; sz = ks_target(sz), nodes *= 2) {
      const uint64_t t = ks_target(sz);
      per_trial = std::max<size_t>(per_trial, (size_t)(nodes * 2 * (sz * sz + t * t) * 8 + nodes * sz * sz * 8));
    }
  }
  size_t batch = per_trial ? std::max<size_t>(1, (size_t)(0.5 * (double)free_b) / per_trial) : ntrials;
  batch = std::min<size_t>(batch, ntrials);
  rc = CAC_OK;
  for (uint32_t t0 = 0; t0 < ntrials && rc == CAC_OK; t0 += (uint32_t)batch) {
    std::vector<KsRoot> roots;
    for (uint32_t t = t0; t < std::min<uint32_t>(ntrials, t0 + (uint32_t)batch); ++t)
      roots.push_back({d_adj, n, first_trial + t, t, bound});
    rc = ks_run(ctx, roots, seed, d_res, s);
  }
  if (rc == CAC_OK) {
    std::vector<int64_t> res(ntrials);
    CAC_HIP(ctx, hipMemcpyAsync(res.data(), d_res, 8 * (size_t)ntrials, hipMemcpyDeviceToHost, s));
    CAC_HIP(ctx, hipStreamSynchronize(s));
    int64_t mn = INT64_MAX;
    for (uint32_t t = 0; t < ntrials; ++t) {
      if (raw_cuts) raw_cuts[t] = res[t];
      mn = std::min(mn, res[t]);
    }
    if (min_cut) *min_cut = mn;
  }
  CAC_HIP(ctx, hipStreamSynchronize(s));
  (void)hipFree(d_bound);
  (void)hipFree(d_res);
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
