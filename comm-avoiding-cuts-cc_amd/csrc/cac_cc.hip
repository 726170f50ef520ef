// cac_cc.hip -- connected components on gfx950: lock-free union-find with
// atomicCAS hooking of the larger root under the smaller, path splitting on every
// find, and a final pointer-jumping pass.
//
// Replaces UnweightedIteratedSparseSampling::connectedComponents
// (src/UnweightedIteratedSparseSampling.cpp:40-72) and its root-sequential
// DisjointSets union loop (src/IteratedSparseSampling.hpp:182-224).  Because a
// root is only ever hooked under a smaller root, parent[x] <= x always holds and
// each tree's root is the smallest vertex id of its component: the labels are
// canonical by construction, whatever order the hooks land in.
//
// Pipeline (edge list resident in HBM, labels array doubles as the parent array):
//   1. init      parent[v] = v                                    (4n bytes written)
//   2. runs      hook the first and the last edge of every run of equal u in
//                the edge list (the whole list is read, ~8% of it hooked on a
//                u-sorted R-MAT list); this builds the giant component
//   3. settle    pointer-jump every vertex to its root
//   4. mode      the most frequent root among 4096 sampled vertices: the giant
//   5. bitmap    bit v = (root(v) == giant root); n/8 bytes, small enough to
//                stay resident in each XCD's L2
//   6. hook      stream every edge; skip it when both endpoint bits are set
//                (already connected), else union-find link.  The random
//                lookups of almost every edge hit the L2-resident bitmap
//                instead of the n*4-byte parent array
//   7. compress  parent[v] = root(v), count roots               (8n bytes)
//   8. (rank)    dense rank of roots by id, relabel              (scan)
// Steps 2-5 run when m is large; small inputs go straight to a plain hook.
// Any edge order is correct; the run sampling pays off when equal u are
// adjacent (the order of the edge-list files and of the R-MAT generator).
#include <cstring>
#include <mutex>

#include <rocprim/rocprim.hpp>

#include "cac_common.h"
#include "cac_internal.h"
#include "cac_scan.h"

namespace cac {
namespace {

constexpr int kCcBlock = 256;
constexpr int kEdgesPerThread = 8;                   // edges each thread keeps in flight
constexpr int kTileEdges = 64 * kEdgesPerThread;     // edges per wavefront tile in k_cc_runs
constexpr uint64_t kSampleMinEdges = 1ull << 18;     // below this: plain hook only
constexpr int kModeSamples = 4096;
constexpr int kModeTable = 8192;                     // open-addressing table, > kModeSamples
constexpr size_t kCcTmpHead = 64;                    // u32 words: [0] component count, [1] giant root

// Parent-array loads inside the union-find kernels are relaxed atomics at
// workgroup scope: plain cached loads.  A stale parent (an older state of the
// forest, from this XCD's L2) is still an ancestor, so finds stay correct, and
// only the CAS in uf_link has to act on the current value.
//
// Stores are relaxed atomics at agent scope (sc1).  Workgroup-scope stores
// (sc0) were measured on gfx950 to lose later writes to the same word: a
// path-splitting store left in one XCD's L2 can land in memory after the next
// kernels rewrote that word from another XCD, so compressed labels reverted to
// a former root.  Agent-scope stores cost nothing measurable here.
__device__ inline uint32_t ld_par(const uint32_t* p, uint32_t i) {
  return __hip_atomic_load(p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void st_par(uint32_t* p, uint32_t i, uint32_t v) {
  __hip_atomic_store(p + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Agent-scope load for the passes that read the finished forest of the
// previous kernel (settle, mode, bitmap, compress).
__device__ inline uint32_t ld_par_fresh(const uint32_t* p, uint32_t i) {
  return __hip_atomic_load(p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t uf_find_fresh(uint32_t* par, uint32_t x) {
  uint32_t p = ld_par_fresh(par, x);
  while (p < x) {
    x = p;
    p = ld_par_fresh(par, x);
  }
  return x;
}

// Root of x with path splitting: every visited node is re-pointed at its
// grandparent.  Any ancestor is a valid parent, so racing writers are benign.
__device__ inline uint32_t uf_find_from(uint32_t* par, uint32_t x, uint32_t p /* = par[x] */) {
  while (p < x) {
    uint32_t g = ld_par(par, p);
    if (g < p) st_par(par, x, g);
    x = p;
    p = g;
  }
  return x;
}

__device__ inline uint32_t uf_find(uint32_t* par, uint32_t x) { return uf_find_from(par, x, ld_par(par, x)); }

// Union of the sets of a and b (pa, pb: their already-loaded parents): CAS the
// larger root to point at the smaller.  Returns 1 when this call hooked a root
// (two trees became one), 0 when a and b were already in one tree.
__device__ inline uint32_t uf_link_from(uint32_t* par, uint32_t a, uint32_t pa, uint32_t b, uint32_t pb) {
  uint32_t ra = uf_find_from(par, a, pa), rb = uf_find_from(par, b, pb);
  while (ra != rb) {
    uint32_t lo = ra < rb ? ra : rb;
    uint32_t hi = ra < rb ? rb : ra;
    uint32_t old = atomicCAS(par + hi, hi, lo);
    if (old == hi) return 1u;
    // hi was hooked elsewhere meanwhile: climb from its new parent and retry.
    ra = uf_find(par, old);
    rb = uf_find(par, lo);
  }
  return 0u;
}

// Union without path splitting: the only stores are CAS hooks of roots.  Used
// once vertices have been re-pointed at their final roots (k_af_final,
// k_af_heavy): a path-splitting store built from a stale (older) parent could
// otherwise put back an older ancestor over such a vertex's root.
__device__ inline uint32_t uf_find_ro(const uint32_t* par, uint32_t x) {
  uint32_t p = ld_par(par, x);
  while (p < x) {
    x = p;
    p = ld_par(par, x);
  }
  return x;
}

__device__ inline uint32_t uf_link_ro(uint32_t* par, uint32_t a, uint32_t b) {
  uint32_t ra = uf_find_ro(par, a), rb = uf_find_ro(par, b);
  while (ra != rb) {
    const uint32_t lo = ra < rb ? ra : rb, hi = ra < rb ? rb : ra;
    const uint32_t old = atomicCAS(par + hi, hi, lo);
    if (old == hi) return 1u;
    ra = uf_find_ro(par, old);
    rb = uf_find_ro(par, lo);
  }
  return 0u;
}

__device__ inline uint32_t uf_link(uint32_t* par, uint32_t a, uint32_t b) {
  return uf_link_from(par, a, ld_par(par, a), b, ld_par(par, b));
}

__global__ void __launch_bounds__(kCcBlock) k_cc_init(uint32_t* __restrict__ par, uint32_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * kCcBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kCcBlock)
    par[i] = (uint32_t)i;
}

__device__ inline uint64_t ld_edge_nt(const cac_uedge* e, uint64_t i) {
  return __builtin_nontemporal_load((const unsigned long long*)e + i);
}

// Wave-tiled: wavefront w handles tiles w, w + W, ...; edge k * 64 + lane of a
// tile sits in register k of that lane, so every load instruction is coalesced
// and the neighbours of an edge (for the run test) are one lane away.
__global__ void __launch_bounds__(kCcBlock) k_cc_runs(const cac_uedge* __restrict__ edges, uint64_t m,
                                                      uint32_t* par) {
  const int lane = threadIdx.x & 63;
  const uint64_t waves = (uint64_t)gridDim.x * (kCcBlock / 64);
  const uint64_t tiles = (m + kTileEdges - 1) / kTileEdges;
  for (uint64_t t = (uint64_t)blockIdx.x * (kCcBlock / 64) + (threadIdx.x >> 6); t < tiles; t += waves) {
    const uint64_t base = t * kTileEdges;
    uint32_t u[kEdgesPerThread], v[kEdgesPerThread];
#pragma unroll
    for (int k = 0; k < kEdgesPerThread; ++k) {
      const uint64_t i = base + (uint64_t)k * 64 + lane;
      const uint64_t e = i < m ? ld_edge_nt(edges, i) : ~0ull;  // ~0: past the end, never equal to an id
      u[k] = (uint32_t)e;
      v[k] = (uint32_t)(e >> 32);
    }
    const uint32_t before = base > 0 ? edges[base - 1].u : ~0u;
    const uint32_t after = base + kTileEdges < m ? edges[base + kTileEdges].u : ~0u;
#pragma unroll
    for (int k = 0; k < kEdgesPerThread; ++k) {
      uint32_t up = __shfl_up(u[k], 1, 64);
      uint32_t dn = __shfl_down(u[k], 1, 64);
      const uint32_t prev_last = __shfl(u[k > 0 ? k - 1 : 0], 63, 64);
      const uint32_t next_first = __shfl(u[k + 1 < kEdgesPerThread ? k + 1 : k], 0, 64);
      if (lane == 0) up = k > 0 ? prev_last : before;
      if (lane == 63) dn = k + 1 < kEdgesPerThread ? next_first : after;
      const bool valid = base + (uint64_t)k * 64 + lane < m;
      if (valid && u[k] != v[k] && (u[k] != up || u[k] != dn)) uf_link(par, u[k], v[k]);
    }
  }
}

// Pointer jumping over [0, n): every vertex is re-pointed at its root; returns
// the number of roots this thread saw.
__device__ inline uint32_t jump_pass(uint32_t* par, uint32_t n) {
  uint32_t roots = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * kCcBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kCcBlock) {
    const uint32_t v = (uint32_t)i;
    const uint32_t p = ld_par_fresh(par, v);
    if (p == v) {
      ++roots;
      continue;
    }
    const uint32_t r = uf_find_fresh(par, p);
    if (r != p) st_par(par, v, r);
  }
  return roots;
}

__global__ void __launch_bounds__(kCcBlock) k_cc_settle(uint32_t* par, uint32_t n) { (void)jump_pass(par, n); }

// Most frequent root among kModeSamples vertices spread over [0, n), each
// sample climbing to its root (one hop on a settled forest).  Only a
// performance choice: any root gives correct labels.
__global__ void __launch_bounds__(1024) k_cc_mode(const uint32_t* __restrict__ par, uint32_t n,
                                                  uint32_t* __restrict__ out_root) {
  __shared__ uint32_t key[kModeTable], cnt[kModeTable];
  __shared__ uint64_t best_w[1024 / 64];
  for (int i = threadIdx.x; i < kModeTable; i += 1024) {
    key[i] = ~0u;
    cnt[i] = 0;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < kModeSamples; j += 1024) {
    uint32_t h = (uint32_t)j * 0x9E3779B1u;  // scatter the sample positions
    h ^= h >> 15;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    uint32_t r = (uint32_t)(((uint64_t)h * n) >> 32), up = ld_par_fresh(par, r);
    while (up < r) {
      r = up;
      up = ld_par_fresh(par, r);
    }
    uint32_t slot = (r * 0x9E3779B1u) >> 19;  // 13 bits = kModeTable slots
    for (;;) {
      const uint32_t old = atomicCAS(&key[slot], ~0u, r);
      if (old == ~0u || old == r) {
        atomicAdd(&cnt[slot], 1u);
        break;
      }
      slot = (slot + 1) & (kModeTable - 1);
    }
  }
  __syncthreads();
  uint64_t best = 0;  // (count << 32) | ~root: larger count wins, then smaller root
  for (int i = threadIdx.x; i < kModeTable; i += 1024)
    if (cnt[i]) best = max(best, ((uint64_t)cnt[i] << 32) | (uint64_t)(~key[i]));
  for (int d = 32; d > 0; d >>= 1) best = max(best, (uint64_t)__shfl_down(best, d, 64));
  if ((threadIdx.x & 63) == 0) best_w[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 1024 / 64; ++w) best = max(best, best_w[w]);
    out_root[0] = ~(uint32_t)best;
  }
}

// bits[v / 64] bit (v % 64) = (par[v] == root); one ballot per 64 vertices.
__global__ void __launch_bounds__(kCcBlock) k_cc_bitmap(const uint32_t* __restrict__ par, uint32_t n,
                                                        const uint32_t* __restrict__ root, uint64_t* __restrict__ bits) {
  const uint32_t R = root[0];
  const int lane = threadIdx.x & 63;
  const uint64_t waves = (uint64_t)gridDim.x * (kCcBlock / 64);
  for (uint64_t w = (uint64_t)blockIdx.x * (kCcBlock / 64) + (threadIdx.x >> 6); w * 64 < n; w += waves) {
    const uint64_t v = w * 64 + lane;
    const bool in = v < n && ld_par_fresh(par, (uint32_t)v) == R;
    const uint64_t mask = __ballot(in);
    if (lane == 0) bits[w] = mask;
  }
}

__device__ inline bool bit_of(const uint32_t* bits, uint32_t x) { return (bits[x >> 5] >> (x & 31)) & 1u; }

// Thread t of the grid owns edges t, t + T, t + 2T, ... (T = total threads) so
// each load instruction is coalesced; kEdgesPerThread loads are issued before
// any is consumed.  With a bitmap, an edge whose two endpoints are both in the
// giant component is dropped before any parent lookup.
template <bool kBitmap>
__global__ void __launch_bounds__(kCcBlock) k_cc_hook(const cac_uedge* __restrict__ edges, uint64_t m,
                                                      uint32_t* par, const uint32_t* __restrict__ bits) {
  const uint64_t T = (uint64_t)gridDim.x * kCcBlock;
  uint64_t i0 = (uint64_t)blockIdx.x * kCcBlock + threadIdx.x;
  for (; i0 < m; i0 += T * kEdgesPerThread) {
    uint64_t e[kEdgesPerThread];
#pragma unroll
    for (int k = 0; k < kEdgesPerThread; ++k) {
      uint64_t i = i0 + (uint64_t)k * T;
      e[k] = i < m ? ld_edge_nt(edges, i) : 0ull;  // empty slot = loop (0, 0)
    }
    bool live[kEdgesPerThread];
#pragma unroll
    for (int k = 0; k < kEdgesPerThread; ++k) {
      const uint32_t a = (uint32_t)e[k], b = (uint32_t)(e[k] >> 32);
      live[k] = a != b;
      if (kBitmap) live[k] = live[k] && !(bit_of(bits, a) && bit_of(bits, b));
    }
    // all parent lookups of the batch are in flight before any is consumed
    uint32_t pu[kEdgesPerThread], pv[kEdgesPerThread];
#pragma unroll
    for (int k = 0; k < kEdgesPerThread; ++k) {
      pu[k] = live[k] ? ld_par(par, (uint32_t)e[k]) : 0u;
      pv[k] = live[k] ? ld_par(par, (uint32_t)(e[k] >> 32)) : 0u;
    }
#pragma unroll
    for (int k = 0; k < kEdgesPerThread; ++k) {
      if (!live[k] || pu[k] == pv[k]) continue;
      uf_link_from(par, (uint32_t)e[k], pu[k], (uint32_t)(e[k] >> 32), pv[k]);
    }
  }
}

// Final pointer jumping; counts roots (components).
__global__ void __launch_bounds__(kCcBlock) k_cc_compress(uint32_t* par, uint32_t n, uint32_t* ncomp) {
  uint32_t roots = jump_pass(par, n);
  // block reduction, then one atomic per block: same-address atomics serialise,
  // so one per wave (2^18 of them at n = 2^24) would dominate this kernel
  __shared__ uint32_t wave_sum[kCcBlock / 64];
  for (int d = 32; d > 0; d >>= 1) roots += __shfl_down(roots, d, 64);
  if ((threadIdx.x & 63) == 0) wave_sum[threadIdx.x >> 6] = roots;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < kCcBlock / 64; ++w) t += wave_sum[w];
    if (t) atomicAdd(ncomp, t);
  }
}

__global__ void k_zero_u32(uint32_t* p, int count) {
  if ((int)threadIdx.x < count) p[threadIdx.x] = 0;
}

struct RootFlag {
  const uint32_t* lab;
  __device__ uint32_t operator()(uint64_t i) const { return lab[i] == (uint32_t)i ? 1u : 0u; }
};
struct RootRankStore {
  uint32_t* rank_of;
  __device__ void operator()(uint64_t i, uint32_t ex, uint32_t flag) const {
    if (flag) rank_of[i] = ex;
  }
};

__global__ void __launch_bounds__(kCcBlock) k_cc_apply_rank(uint32_t* lab, const uint32_t* rank_of, uint32_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * kCcBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kCcBlock)
    lab[i] = rank_of[lab[i]];
}

__global__ void __launch_bounds__(kCcBlock) k_check_ids(const cac_uedge* e, uint64_t m, uint32_t n, uint32_t* bad) {
  uint32_t b = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * kCcBlock + threadIdx.x; i < m; i += (uint64_t)gridDim.x * kCcBlock) {
    cac_uedge x = e[i];
    if (x.u >= n || x.v >= n) b = 1;
  }
  if (b) atomicOr(bad, 1u);
}

// ---- connected components on a symmetric CSR (Afforest) ---------------------
// Sutton, Ben-Nun, Barak, "Optimizing Parallel Graph Connectivity Computation via
// Subgraph Sampling" (IPDPS 2018): link a sample of edges, find the most frequent
// root, then link every remaining edge of the vertices outside that component.
// Every edge is in both adjacency lists, so an edge from a vertex of the big
// component to one outside it is still linked from the outside end.  The
// sample: the low-id hub ranges link their first two neighbours, every other
// vertex its first (smallest) one; so the final pass starts at index kAfSkip.
// Linking the second neighbour in the bulk pass as well cost more (its parent
// lookups are random lines of the parent array) than the final pass saves.
constexpr uint32_t kAfSkip = 1;
constexpr uint32_t kAfHeavy = 1024;  // a vertex with more remaining neighbours goes to a workgroup
// Successful hooks are counted per wavefront and added to one of kHookSlots
// counters 64 bytes apart, so no single word takes an atomic from every wave.
// Components = n - hooks: each hook joins two trees into one.
constexpr uint32_t kHookSlots = 64, kHookStride = 16;
// CSR scratch head: [0] components, [1] sampled root, [2] heavy-queue length,
// then the hook slots.
constexpr size_t kAfHead = 64 + (size_t)kHookSlots * kHookStride;

// Per wavefront, without a barrier: a wave with hooks adds them to the slot of
// its global wave index.  (A block barrier here would hold every wave of a
// block until its slowest one, in the divergent final pass, had finished.)
__device__ inline void add_hooks(uint32_t* slots, uint32_t h) {
  for (int d = 32; d > 0; d >>= 1) h += __shfl_down(h, d, 64);
  const uint32_t wave = blockIdx.x * (kCcBlock / 64) + (threadIdx.x >> 6);
  if ((threadIdx.x & 63) == 0 && h) atomicAdd(slots + (wave % kHookSlots) * kHookStride, h);
}

// parent[v] = v with 16-byte stores where the array is aligned, and the scratch
// head (counters and hook slots) cleared.
__global__ void __launch_bounds__(kCcBlock) k_af_init(uint32_t* __restrict__ par, uint32_t n,
                                                      uint32_t* __restrict__ head, uint32_t head_words) {
  const uint64_t T = (uint64_t)gridDim.x * kCcBlock, tid = (uint64_t)blockIdx.x * kCcBlock + threadIdx.x;
  if (tid < head_words) head[tid] = 0;
  uint64_t tail = 0;
  if ((reinterpret_cast<uintptr_t>(par) & 15) == 0) {
    uint4* p4 = reinterpret_cast<uint4*>(par);
    for (uint64_t q = tid; q < n / 4; q += T) {
      const uint32_t v = (uint32_t)(q * 4);
      p4[q] = make_uint4(v, v + 1, v + 2, v + 3);
    }
    tail = (uint64_t)(n / 4) * 4;
  }
  for (uint64_t i = tail + tid; i < n; i += T) par[i] = (uint32_t)i;
}

// Root of the largest component, guessed from kAfSamples vertices of [0, range)
// by a Boyer-Moore majority vote: per lane over its samples, then across the
// wavefront (pairwise merge; ties keep the smaller id, so both partners of a
// butterfly step agree).  Only a performance choice: labels are correct
// whatever root is returned.  One wavefront; the samples' loads are issued
// four at a time (this runs inside the bulk pass, whose register budget, and
// so occupancy, it must not raise).
constexpr int kAfSamples = 1024;
__device__ inline void af_majority_root(uint32_t* par, uint32_t range, uint32_t* out) {
  constexpr int kPer = kAfSamples / 64, kBatch = 4;
  const int lane = threadIdx.x & 63;
  uint32_t cand = ~0u, cnt = 0;
  for (int t0 = 0; t0 < kPer; t0 += kBatch) {
    uint32_t x[kBatch], p[kBatch];
#pragma unroll
    for (int t = 0; t < kBatch; ++t) {
      uint32_t h = (uint32_t)((t0 + t) * 64 + lane) * 0x9E3779B1u;  // scatter the sample positions
      h ^= h >> 15;
      h *= 0x85EBCA6Bu;
      h ^= h >> 13;
      x[t] = (uint32_t)(((uint64_t)h * range) >> 32);
      p[t] = ld_par_fresh(par, x[t]);
    }
#pragma unroll
    for (int t = 0; t < kBatch; ++t) {
      const uint32_t r = p[t] == x[t] ? x[t] : uf_find_fresh(par, p[t]);
      if (cnt == 0) {
        cand = r;
        cnt = 1;
      } else {
        cnt = cand == r ? cnt + 1 : cnt - 1;
      }
    }
  }
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t oc = __shfl_xor(cand, d, 64), on = __shfl_xor(cnt, d, 64);
    if (oc == cand) {
      cnt += on;
    } else if (on > cnt || (on == cnt && oc < cand)) {
      cand = oc;
      cnt = on - cnt;
    } else {
      cnt -= on;
    }
  }
  if (lane == 0) out[0] = cand;
}

// Vertices [v_lo, v_hi) link with their first neighbours (lists are sorted by
// id, so these tend to be the low-id hubs of skewed graphs): two in the hub
// ranges, one in the bulk range.  kBulk: the last range (every vertex above the
// hub prefixes [0, v_lo), which are linked and pointer-jumped by now).  Its
// first wavefront also guesses the root of the largest component from the hub
// forest (af_majority_root), which saves a launch before k_af_final.
template <bool kBulk>
__global__ void __launch_bounds__(kCcBlock) k_af_round(const uint32_t* __restrict__ off,
                                                       const uint32_t* __restrict__ cols, uint32_t v_lo, uint32_t v_hi,
                                                       uint32_t* par, uint32_t* hook_slots, uint32_t* root_out) {
  if (kBulk && blockIdx.x == 0 && threadIdx.x < 64) af_majority_root(par, v_lo, root_out);
  uint32_t hooks = 0;
  for (uint64_t i = v_lo + (uint64_t)blockIdx.x * kCcBlock + threadIdx.x; i < v_hi; i += (uint64_t)gridDim.x * kCcBlock) {
    const uint32_t u = (uint32_t)i, b = off[u], e = off[u + 1];
    uint32_t v0 = u, v1 = u;
    if (kBulk) {
      if (b < e) v0 = cols[b];
    } else if (b + 1 < e) {  // both first neighbours with one 8-byte load (4-byte aligned is enough)
      uint2 x;
      __builtin_memcpy(&x, cols + b, 8);
      v0 = x.x;
      v1 = x.y;
    } else if (b < e) {
      v0 = cols[b];
    }
    // u is taken to be a root without loading par[u]: the CAS checks it, and a
    // failed CAS climbs from the value it returns
    if (v0 != u) hooks += uf_link_from(par, u, u, v0, ld_par(par, v0));
    if (!kBulk && v1 != u) hooks += uf_link(par, u, v1);
  }
  add_hooks(hook_slots, hooks);
}

// Every vertex is re-pointed at its root (the pointer jumping of jump_pass,
// fused here); those whose root is not the sampled root R then link their
// neighbours from index kAfSkip on.  A wavefront links the remaining
// neighbours of all its vertices together: an exclusive scan of the per-lane
// counts lays them out as one list, and lane j of each 64-entry step links
// entry j with its owner (found by a binary search over the scan), so one long
// list does not leave 63 lanes idle.  Lists longer than kAfHeavy are queued for
// k_af_heavy.  A root found here may be hooked under a smaller one by a
// concurrent link: it stays an ancestor, so the store is still valid, and the
// vertex at worst links neighbours that were already joined.  Such a vertex
// may then hold a stale root, so bit v of `outside` marks every vertex with
// neighbours whose root was not R; k_af_fix re-points those.  Offsets are read
// only for vertices outside R's tree.
constexpr int kAfChunks = 4;  // 64-vertex chunks a wavefront of k_af_final has in flight
__global__ void __launch_bounds__(kCcBlock) k_af_final(const uint32_t* __restrict__ off,
                                                       const uint32_t* __restrict__ cols, uint32_t n, uint32_t* par,
                                                       const uint32_t* __restrict__ root, uint32_t* heavy,
                                                       uint32_t* heavy_count, uint64_t* __restrict__ outside,
                                                       uint32_t* hook_slots) {
  constexpr int K = kAfChunks;
  const uint32_t R = root[0];
  const int lane = threadIdx.x & 63;
  uint32_t hooks = 0;
  // Whole wavefronts step together (the ballot and the shuffles see all 64
  // lanes); each step takes K chunks and issues every load of a phase for all of
  // them before using any: the parents, the parents' parents, the offsets.
  const uint64_t wave = (uint64_t)blockIdx.x * (kCcBlock / 64) + (threadIdx.x >> 6);
  for (uint64_t w0 = wave * 64 * K; w0 < n; w0 += (uint64_t)gridDim.x * (kCcBlock / 64) * 64 * K) {
    uint32_t p[K], g[K], b[K], e[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint64_t i = w0 + (uint64_t)k * 64 + lane;
      p[k] = i < n ? ld_par_fresh(par, (uint32_t)i) : (uint32_t)i;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t u = (uint32_t)(w0 + (uint64_t)k * 64 + lane);
      g[k] = p[k] != u ? ld_par_fresh(par, p[k]) : p[k];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint64_t i = w0 + (uint64_t)k * 64 + lane;
      const uint32_t u = (uint32_t)i;
      const uint32_t r = g[k] == p[k] ? p[k] : uf_find_fresh(par, g[k]);
      if (r != p[k]) st_par(par, u, r);
      const bool out = i < n && r != R;
      b[k] = out ? off[u] : 0u;
      e[k] = out ? off[u + 1] : 0u;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint64_t c0 = w0 + (uint64_t)k * 64;
      const uint64_t mask = __ballot(e[k] > b[k]);
      if (c0 < n && lane == 0) outside[c0 >> 6] = mask;
      if (!mask) continue;
      const uint32_t u = (uint32_t)(c0 + lane), bb = b[k] + kAfSkip;
      uint32_t c = bb < e[k] ? e[k] - bb : 0u;
      if (c > kAfHeavy) {
        heavy[atomicAdd(heavy_count, 1u)] = u;
        c = 0;
      }
      uint32_t incl = c;
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
      }
      const uint32_t total = __shfl(incl, 63, 64);
      for (uint32_t j0 = 0; j0 < total; j0 += 64) {
        const uint32_t j = j0 + lane;
        uint32_t L = 0;  // owner lane: the number of lanes whose inclusive count is <= j
        for (int st = 32; st > 0; st >>= 1)
          if (__shfl(incl, (int)L + st - 1, 64) <= j) L += st;
        const uint32_t ou = __shfl(u, (int)L, 64), ob = __shfl(bb, (int)L, 64);
        const uint32_t ofirst = __shfl(incl - c, (int)L, 64);
        if (j < total) hooks += uf_link_ro(par, ou, cols[ob + (j - ofirst)]);
      }
    }
  }
  add_hooks(hook_slots, hooks);
}

// One workgroup per queued vertex (grid-stride over the queue).
__global__ void __launch_bounds__(kCcBlock) k_af_heavy(const uint32_t* __restrict__ off,
                                                       const uint32_t* __restrict__ cols, uint32_t* par,
                                                       const uint32_t* __restrict__ heavy,
                                                       const uint32_t* __restrict__ heavy_count, uint32_t* hook_slots) {
  const uint32_t count = heavy_count[0];
  uint32_t hooks = 0;
  for (uint32_t q = blockIdx.x; q < count; q += gridDim.x) {
    const uint32_t u = heavy[q];
    const uint32_t b = off[u] + kAfSkip, e = off[u + 1];
    for (uint32_t k = b + threadIdx.x; k < e; k += kCcBlock) hooks += uf_link_ro(par, u, cols[k]);
  }
  add_hooks(hook_slots, hooks);
}

// Final labels.  A vertex whose root was R when k_af_final re-pointed it still
// has the right label unless R itself was hooked later; a vertex without
// neighbours never changes.  So only the vertices marked in `outside` are
// re-pointed at their roots (one bitmap word per lane, its set bits in turn),
// unless R is no longer a root, in which case every vertex is (one vertex per
// lane).  Block 0 also writes the component count, n - hooks.
__global__ void __launch_bounds__(kCcBlock) k_af_fix(uint32_t* par, uint32_t n, const uint64_t* __restrict__ outside,
                                                     const uint32_t* __restrict__ root, uint32_t* counters,
                                                     const uint32_t* hook_slots) {
  const uint32_t R = root[0];
  const uint64_t T = (uint64_t)gridDim.x * kCcBlock, tid = (uint64_t)blockIdx.x * kCcBlock + threadIdx.x;
  if (ld_par_fresh(par, R) != R) {
    for (uint64_t i = tid; i < n; i += T) {
      const uint32_t u = (uint32_t)i, p = ld_par_fresh(par, u);
      if (p == u) continue;
      const uint32_t r = uf_find_fresh(par, p);
      if (r != p) st_par(par, u, r);
    }
  } else {
    for (uint64_t w = tid; w < ((uint64_t)n + 63) / 64; w += T) {
      for (uint64_t bits = outside[w]; bits; bits &= bits - 1) {
        const uint32_t u = (uint32_t)(w * 64 + __builtin_ctzll(bits)), p = ld_par_fresh(par, u);
        if (p == u) continue;
        const uint32_t r = uf_find_fresh(par, p);
        if (r != p) st_par(par, u, r);
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < 64) {
    uint32_t h = ld_par_fresh(hook_slots, threadIdx.x * kHookStride);
    for (int d = 32; d > 0; d >>= 1) h += __shfl_down(h, d, 64);
    if (threadIdx.x == 0) counters[0] = n - h;
  }
}

// ---- symmetric CSR from an edge list ------------------------------------------
// Directed entries (src << bits) | dst of both orientations (loops dropped),
// radix-sorted, give every adjacency list sorted by neighbour id.
__global__ void __launch_bounds__(256) k_csr_keys(const cac_uedge* __restrict__ e, uint64_t m, uint32_t bits,
                                                  uint64_t* __restrict__ keys) {
  const uint64_t loop = ~0ull;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (uint64_t)gridDim.x * 256) {
    const cac_uedge x = e[i];
    const bool lp = x.u == x.v;
    keys[2 * i] = lp ? loop : ((uint64_t)x.u << bits) | x.v;
    keys[2 * i + 1] = lp ? loop : ((uint64_t)x.v << bits) | x.u;
  }
}

// cols[i] = dst of sorted entry i; off[v] = first entry with src >= v, for every
// v in (src(i-1), src(i)], and off[v] = nnz past the last source.
__global__ void __launch_bounds__(256) k_csr_split(const uint64_t* __restrict__ keys, uint64_t nnz, uint32_t n,
                                                   uint32_t bits, uint32_t* __restrict__ off, uint32_t* __restrict__ cols) {
  const uint64_t mask = (1ull << bits) - 1;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i <= nnz; i += (uint64_t)gridDim.x * 256) {
    const uint64_t src = i < nnz ? keys[i] >> bits : n;
    const uint64_t prev = i > 0 ? (keys[i - 1] >> bits) : ~0ull;
    if (i < nnz) cols[i] = (uint32_t)(keys[i] & mask);
    if (i == 0) {
      for (uint64_t v = 0; v <= src; ++v) off[v] = 0;
    } else if (src != prev) {
      for (uint64_t v = prev + 1; v <= src; ++v) off[v] = (uint32_t)i;
    }
  }
}

// ---- approximate min cut (ApproximateCut::run) -------------------------------
// (1 - 2^-i)^w in 0.64 fixed point, floor(a*b/2^64) per multiply; ~0 stands for 1.
__device__ inline uint64_t keep_threshold(uint32_t i, uint64_t w) {
  if (w == 0) return ~0ull;
  uint64_t base = (i >= 64) ? ~0ull : 0ull - (1ull << (64 - i));
  bool have = false;
  uint64_t acc = 0;
  while (w) {
    if (w & 1) {
      acc = have ? __umul64hi(acc, base) : base;
      have = true;
    }
    w >>= 1;
    if (w) base = __umul64hi(base, base);
  }
  return acc;
}

// Edge e is kept in copy j at level i iff lo64(Philox(i, j, e)) >= threshold;
// kept edges are hooked into copy j's slice [j n, (j + 1) n) of the forest.
__global__ void __launch_bounds__(kCcBlock) k_approx_hook(const cac_edge* __restrict__ edges, uint64_t m, uint32_t n,
                                                          uint32_t copies, uint32_t level, uint32_t seed,
                                                          uint32_t* par) {
  for (uint64_t e = (uint64_t)blockIdx.x * kCcBlock + threadIdx.x; e < m; e += (uint64_t)gridDim.x * kCcBlock) {
    const cac_edge x = edges[e];
    if (x.w == 0 || x.u == x.v) continue;
    const uint64_t q = keep_threshold(level, x.w);
    for (uint32_t j = 0; j < copies; ++j) {
      const u32x4 r = philox4x32(level, j, (uint32_t)e, (uint32_t)(e >> 32), seed, kDomApprox);
      if (lo64(r) < q) continue;
      uf_link(par, x.u + j * n, x.v + j * n);
    }
  }
}

}  // namespace

size_t cc_bitmap_words(uint64_t m, uint32_t n) {
  return m >= kSampleMinEdges ? (((size_t)n + 63) / 64 * 2 + 63) & ~size_t(63) : 0;
}
size_t cc_rank_tmp_words(uint32_t n) { return ((size_t)(n + 63) & ~size_t(63)) + scan_tiles(n) + 64; }

// Device scratch of cc_launch, in u32 words: head, bitmap (large m), rank tables
// (label_mode == RANK).
size_t cc_tmp_words(uint64_t m, uint32_t n, int label_mode) {
  return kCcTmpHead + cc_bitmap_words(m, n) + (label_mode == CAC_LABEL_RANK ? cc_rank_tmp_words(n) : 0);
}

// Launch sequence shared by cac_cc and cac_cc_dev; d_tmp holds cc_tmp_words()
// u32 words, 256-byte aligned.  The component count lands in d_tmp[0].
int cc_launch(cac_ctx* ctx, const cac_uedge* d_edges, uint64_t m, uint32_t n, int label_mode, uint32_t* d_lab,
              uint32_t* d_tmp, hipStream_t s) {
  uint32_t* d_counter = d_tmp;
  uint32_t* d_rank_tmp = d_tmp + kCcTmpHead + cc_bitmap_words(m, n);
  const unsigned cap = 256 * 64;
  prof_begin(ctx, "cc_init", s);
  hipLaunchKernelGGL(k_cc_init, dim3(grid_for(n, kCcBlock, cap)), dim3(kCcBlock), 0, s, d_lab, n);
  prof_end(ctx, s);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, d_counter, 1);
  const bool use_sample = m >= kSampleMinEdges;
  if (use_sample) {
    uint32_t* d_root = d_counter + 1;
    uint64_t* d_bits = (uint64_t*)(d_counter + kCcTmpHead);
    prof_begin(ctx, "cc_runs", s);
    unsigned gr = grid_for((m + kTileEdges - 1) / kTileEdges, kCcBlock / 64, 256 * 32);
    hipLaunchKernelGGL(k_cc_runs, dim3(gr), dim3(kCcBlock), 0, s, d_edges, m, d_lab);
    prof_end(ctx, s);
    prof_begin(ctx, "cc_settle", s);
    hipLaunchKernelGGL(k_cc_settle, dim3(grid_for(n, kCcBlock, cap)), dim3(kCcBlock), 0, s, d_lab, n);
    prof_end(ctx, s);
    prof_begin(ctx, "cc_bitmap", s);
    hipLaunchKernelGGL(k_cc_mode, dim3(1), dim3(1024), 0, s, (const uint32_t*)d_lab, n, d_root);
    hipLaunchKernelGGL(k_cc_bitmap, dim3(grid_for(((uint64_t)n + 63) / 64, kCcBlock / 64, cap)), dim3(kCcBlock), 0, s,
                       (const uint32_t*)d_lab, n, (const uint32_t*)d_root, d_bits);
    prof_end(ctx, s);
    prof_begin(ctx, "cc_hook", s);
    unsigned g = grid_for((m + kEdgesPerThread - 1) / kEdgesPerThread, kCcBlock, 256 * 32);
    hipLaunchKernelGGL(k_cc_hook<true>, dim3(g), dim3(kCcBlock), 0, s, d_edges, m, d_lab, (const uint32_t*)d_bits);
    prof_end(ctx, s);
  } else if (m) {
    prof_begin(ctx, "cc_hook", s);
    unsigned g = grid_for((m + kEdgesPerThread - 1) / kEdgesPerThread, kCcBlock, 256 * 32);
    hipLaunchKernelGGL(k_cc_hook<false>, dim3(g), dim3(kCcBlock), 0, s, d_edges, m, d_lab, (const uint32_t*)nullptr);
    prof_end(ctx, s);
  }
  prof_begin(ctx, "cc_compress", s);
  hipLaunchKernelGGL(k_cc_compress, dim3(grid_for(n, kCcBlock, 4096)), dim3(kCcBlock), 0, s, d_lab, n, d_counter);
  prof_end(ctx, s);
  if (label_mode == CAC_LABEL_RANK) {
    uint32_t* rank_of = d_rank_tmp;
    uint32_t* partial = d_rank_tmp + ((n + 63) & ~63u);
    prof_begin(ctx, "cc_rank", s);
    hipError_t e = device_exclusive_scan<uint32_t>(RootFlag{d_lab}, RootRankStore{rank_of}, n, partial,
                                                   (uint32_t*)nullptr, s);
    if (e != hipSuccess) return hip_fail(ctx, e, "rank scan");
    hipLaunchKernelGGL(k_cc_apply_rank, dim3(grid_for(n, kCcBlock, cap)), dim3(kCcBlock), 0, s, d_lab,
                       (const uint32_t*)rank_of, n);
    prof_end(ctx, s);
  }
  CAC_LAUNCH_CHECK(ctx, "cc kernels");
  return CAC_OK;
}

// Queue of long adjacency lists: n + 64 words, rounded to 64 so what follows
// stays 256-byte aligned.
static size_t af_queue_words(uint32_t n) { return ((size_t)n + 64 + 63) & ~size_t(63); }

// Afforest on a symmetric CSR.  Scratch: the head (counters and hook slots), the
// queue, the `outside` bitmap (one bit per vertex, whole u64 words), then the
// rank tables.
size_t cc_csr_tmp_words(uint32_t n, int label_mode) {
  return kAfHead + af_queue_words(n) + 2 * (((size_t)n + 63) / 64) +
         (label_mode == CAC_LABEL_RANK ? cc_rank_tmp_words(n) : 0);
}

int cc_csr_launch(cac_ctx* ctx, const uint32_t* d_off, const uint32_t* d_cols, uint32_t n, int label_mode,
                  uint32_t* d_lab, uint32_t* d_tmp, hipStream_t s) {
  uint32_t* d_counter = d_tmp;  // [0] components, [1] sampled root, [2] queue length
  uint32_t* d_hooks = d_tmp + 64;
  uint32_t* d_heavy = d_tmp + kAfHead;
  uint64_t* d_outside = reinterpret_cast<uint64_t*>(d_heavy + af_queue_words(n));
  uint32_t* d_rank_tmp = reinterpret_cast<uint32_t*>(d_outside + ((size_t)n + 63) / 64);
  const unsigned cap = 256 * 64;
  prof_begin(ctx, "af_init", s);
  {
    unsigned g = grid_for(((uint64_t)n + 3) / 4, kCcBlock, cap);
    const unsigned g_head = (unsigned)((kAfHead + kCcBlock - 1) / kCcBlock);
    hipLaunchKernelGGL(k_af_init, dim3(g > g_head ? g : g_head), dim3(kCcBlock), 0, s, d_lab, n, d_counter,
                       (uint32_t)kAfHead);
  }
  prof_end(ctx, s);
  // Low ids first, each prefix pointer-jumped before the next range links to it:
  // the first neighbours of most vertices are low-id hubs, whose chains are then
  // one hop long (0.99 ms -> 0.58 ms for the round on R-MAT 2^24 / 2^28).
  prof_begin(ctx, "af_hubs", s);
  uint32_t lo = 0;
  for (uint32_t hb : {1u << 12, 1u << 16}) {
    const uint32_t hi = n < hb ? n : hb;
    if (hi <= lo) continue;
    hipLaunchKernelGGL(k_af_round<false>, dim3(grid_for(hi - lo, kCcBlock, cap)), dim3(kCcBlock), 0, s, d_off, d_cols,
                       lo, hi, d_lab, d_hooks, (uint32_t*)nullptr);
    hipLaunchKernelGGL(k_cc_settle, dim3(grid_for(hi, kCcBlock, cap)), dim3(kCcBlock), 0, s, d_lab, hi);
    lo = hi;
  }
  prof_end(ctx, s);
  if (lo < n) {
    prof_begin(ctx, "af_round", s);
    hipLaunchKernelGGL(k_af_round<true>, dim3(grid_for(n - lo, kCcBlock, cap)), dim3(kCcBlock), 0, s, d_off, d_cols, lo,
                       n, d_lab, d_hooks, d_counter + 1);
    prof_end(ctx, s);
  }
  if (lo >= n) {  // no bulk range (n within the hub prefixes): a separate sampling launch
    prof_begin(ctx, "af_mode", s);
    hipLaunchKernelGGL(k_cc_mode, dim3(1), dim3(1024), 0, s, (const uint32_t*)d_lab, n, d_counter + 1);
    prof_end(ctx, s);
  }
  prof_begin(ctx, "af_final", s);
  hipLaunchKernelGGL(k_af_final, dim3(grid_for(((uint64_t)n + 64 * kAfChunks - 1) / (64 * kAfChunks), kCcBlock / 64, 4096)),
                     dim3(kCcBlock), 0, s, d_off, d_cols, n, d_lab,
                     (const uint32_t*)(d_counter + 1), d_heavy, d_counter + 2, d_outside, d_hooks);
  hipLaunchKernelGGL(k_af_heavy, dim3(256), dim3(kCcBlock), 0, s, d_off, d_cols, d_lab, (const uint32_t*)d_heavy,
                     (const uint32_t*)(d_counter + 2), d_hooks);
  prof_end(ctx, s);
  prof_begin(ctx, "af_fix", s);
  hipLaunchKernelGGL(k_af_fix, dim3(grid_for(n, kCcBlock, 1024)), dim3(kCcBlock), 0, s, d_lab, n,
                     (const uint64_t*)d_outside, (const uint32_t*)(d_counter + 1), d_counter,
                     (const uint32_t*)d_hooks);
  prof_end(ctx, s);
  if (label_mode == CAC_LABEL_RANK) {
    uint32_t* rank_of = d_rank_tmp;
    uint32_t* partial = d_rank_tmp + ((n + 63) & ~63u);
    hipError_t e = device_exclusive_scan<uint32_t>(RootFlag{d_lab}, RootRankStore{rank_of}, n, partial,
                                                   (uint32_t*)nullptr, s);
    if (e != hipSuccess) return hip_fail(ctx, e, "rank scan");
    hipLaunchKernelGGL(k_cc_apply_rank, dim3(grid_for(n, kCcBlock, cap)), dim3(kCcBlock), 0, s, d_lab,
                       (const uint32_t*)rank_of, n);
  }
  CAC_LAUNCH_CHECK(ctx, "csr cc kernels");
  return CAC_OK;
}

// ---- connected components without allocation (cac_cc_static_dev) -----------
// The scratch of cc_launch for up to kStaticMaxN vertices lives in one static
// device array: the head (counters), the giant-component bitmap, and for rank
// labels the per-word root prefix and the scan partials.  Rank labels are
// computed from a root bitmap (which reuses the giant bitmap's words, free once
// the hook pass has run) instead of cc_launch's n-entry rank table:
// rank(r) = roots in the words before r's + roots below r in its word.
constexpr uint32_t kStaticMaxN = CAC_CC_STATIC_MAX_N;
constexpr size_t kStaticBitsWords = (size_t)kStaticMaxN / 64 * 2;  // one u64 per 64 vertices
constexpr size_t kStaticPreWords = (size_t)kStaticMaxN / 64;       // one u32 per bitmap word
constexpr size_t kStaticPartWords = ((size_t)kStaticMaxN / 64 + kScanTile - 1) / kScanTile + 64;
constexpr size_t kStaticWords = kCcTmpHead + kStaticBitsWords + kStaticPreWords + kStaticPartWords;
static_assert(kStaticMaxN % 64 == 0, "whole bitmap words");

static __device__ __align__(256) uint32_t cc_static_scratch[kStaticWords];

namespace {

// The array's device address, one per device (each GPU has its own copy),
// taken by hipGetSymbolAddress with that device current.
constexpr int kStaticMaxDevices = 64;
std::mutex g_static_mu;
void* g_static_ptr[kStaticMaxDevices] = {};

// bits[w] bit b = (lab[64 w + b] is a root); one ballot per 64 vertices.
__global__ void __launch_bounds__(kCcBlock) k_cc_root_bits(const uint32_t* __restrict__ lab, uint32_t n,
                                                           uint64_t* __restrict__ bits) {
  const int lane = threadIdx.x & 63;
  const uint64_t waves = (uint64_t)gridDim.x * (kCcBlock / 64);
  for (uint64_t w = (uint64_t)blockIdx.x * (kCcBlock / 64) + (threadIdx.x >> 6); w * 64 < n; w += waves) {
    const uint64_t v = w * 64 + lane;
    const uint64_t mask = __ballot(v < n && lab[v] == (uint32_t)v);
    if (lane == 0) bits[w] = mask;
  }
}

struct WordPop {
  const uint64_t* bits;
  __device__ uint32_t operator()(uint64_t i) const { return (uint32_t)__popcll(bits[i]); }
};
struct WordPreStore {
  uint32_t* pre;
  __device__ void operator()(uint64_t i, uint32_t ex, uint32_t) const { pre[i] = ex; }
};

// lab[v] = rank of its root r among all roots, from the bitmap alone (lab is
// read and written only at v, so the pass has no cross-thread hazard).
__global__ void __launch_bounds__(kCcBlock) k_cc_rank_bits(uint32_t* lab, uint32_t n, const uint64_t* __restrict__ bits,
                                                           const uint32_t* __restrict__ pre) {
  for (uint64_t i = (uint64_t)blockIdx.x * kCcBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kCcBlock) {
    const uint32_t r = lab[i], w = r >> 6;
    lab[i] = pre[w] + (uint32_t)__popcll(bits[w] & ((1ull << (r & 63)) - 1));
  }
}

}  // namespace

int cc_static_init(cac_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kStaticMaxDevices)
    return fail(ctx, CAC_EDEVICE, "static connected-components scratch: device index beyond 63");
  std::lock_guard<std::mutex> lock(g_static_mu);
  if (g_static_ptr[ctx->device]) return CAC_OK;
  void* p = nullptr;
  CAC_HIP(ctx, hipSetDevice(ctx->device));
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(cc_static_scratch));
  if (e != hipSuccess || !p) return hip_fail(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress");
  g_static_ptr[ctx->device] = p;
  return CAC_OK;
}

}  // namespace cac

extern "C" {

int cac_cc_dev(cac_ctx* ctx, const cac_uedge* d_edges, uint64_t m, uint32_t n, int label_mode, uint32_t* d_labels,
               uint32_t* ncomp_out, void* stream) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if ((m && !d_edges) || (n && !d_labels) || !ncomp_out) return cac::fail(ctx, CAC_EINVAL, "null pointer");
  if (label_mode != CAC_LABEL_MIN_ID && label_mode != CAC_LABEL_RANK)
    return cac::fail(ctx, CAC_EINVAL, "label_mode must be CAC_LABEL_MIN_ID or CAC_LABEL_RANK");
  hipStream_t s = cac::pick_stream(ctx, stream);
  size_t words = cac::cc_tmp_words(m, n, label_mode);
  void* base = nullptr;
  rc = cac::scratch_reserve(ctx, words * 4, &base);
  if (rc) return rc;
  uint32_t* counter = (uint32_t*)base;
  rc = cac::cc_launch(ctx, d_edges, m, n, label_mode, d_labels, counter, s);
  if (rc) return rc;
  CAC_HIP(ctx, hipMemcpyAsync(ctx->pinned, counter, 4, hipMemcpyDeviceToHost, s));
  CAC_HIP(ctx, hipStreamSynchronize(s));
  *ncomp_out = *(uint32_t*)ctx->pinned;
  return CAC_OK;
}

int cac_cc_static_dev(cac_ctx* ctx, const cac_uedge* d_edges, uint64_t m, uint32_t n, int label_mode,
                      uint32_t* d_labels, uint32_t* ncomp_out, void* stream) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if ((m && !d_edges) || (n && !d_labels) || !ncomp_out) return cac::fail(ctx, CAC_EINVAL, "null pointer");
  if (label_mode != CAC_LABEL_MIN_ID && label_mode != CAC_LABEL_RANK)
    return cac::fail(ctx, CAC_EINVAL, "label_mode must be CAC_LABEL_MIN_ID or CAC_LABEL_RANK");
  if (n > CAC_CC_STATIC_MAX_N)
    return cac::fail(ctx, CAC_EINVAL, "cac_cc_static_dev: n exceeds CAC_CC_STATIC_MAX_N (2^25 vertices), the size "
                                      "of its static scratch; use cac_cc_dev");
  rc = cac::cc_static_init(ctx);
  if (rc) return rc;
  uint32_t* tmp = (uint32_t*)cac::g_static_ptr[ctx->device];
  hipStream_t s = cac::pick_stream(ctx, stream);
  // per-kernel timing would create events: off for this call
  const bool prof = ctx->profiling;
  ctx->profiling = false;
  rc = cac::cc_launch(ctx, d_edges, m, n, CAC_LABEL_MIN_ID, d_labels, tmp, s);
  if (rc == CAC_OK && label_mode == CAC_LABEL_RANK && n) {
    uint64_t* bits = reinterpret_cast<uint64_t*>(tmp + cac::kCcTmpHead);
    uint32_t* pre = tmp + cac::kCcTmpHead + cac::kStaticBitsWords;
    uint32_t* partial = pre + cac::kStaticPreWords;
    const uint64_t words = ((uint64_t)n + 63) / 64;
    hipLaunchKernelGGL(cac::k_cc_root_bits, dim3(cac::grid_for(words, cac::kCcBlock / 64, 256 * 64)),
                       dim3(cac::kCcBlock), 0, s, (const uint32_t*)d_labels, n, bits);
    hipError_t e = cac::device_exclusive_scan<uint32_t>(cac::WordPop{bits}, cac::WordPreStore{pre}, words, partial,
                                                        (uint32_t*)nullptr, s);
    if (e != hipSuccess) rc = cac::hip_fail(ctx, e, "root prefix scan");
    else
      hipLaunchKernelGGL(cac::k_cc_rank_bits, dim3(cac::grid_for(n, cac::kCcBlock, 256 * 64)), dim3(cac::kCcBlock), 0,
                         s, d_labels, n, (const uint64_t*)bits, (const uint32_t*)pre);
    if (rc == CAC_OK) {
      e = hipGetLastError();
      if (e != hipSuccess) rc = cac::hip_fail(ctx, e, "rank kernels");
    }
  }
  ctx->profiling = prof;
  if (rc) return rc;
  CAC_HIP(ctx, hipMemcpyAsync(ctx->pinned, tmp, 4, hipMemcpyDeviceToHost, s));
  CAC_HIP(ctx, hipStreamSynchronize(s));
  *ncomp_out = *(uint32_t*)ctx->pinned;
  return CAC_OK;
}

int cac_cc(cac_ctx* ctx, const cac_uedge* edges, uint64_t m, uint32_t n, int label_mode, uint32_t* labels_out,
           uint32_t* ncomp_out) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if ((m && !edges) || (n && !labels_out) || !ncomp_out) return cac::fail(ctx, CAC_EINVAL, "null pointer");
  if (label_mode != CAC_LABEL_MIN_ID && label_mode != CAC_LABEL_RANK)
    return cac::fail(ctx, CAC_EINVAL, "label_mode must be CAC_LABEL_MIN_ID or CAC_LABEL_RANK");
  hipStream_t s = ctx->stream;
  size_t e_bytes = cac::carve_size(m * sizeof(cac_uedge));
  size_t l_bytes = cac::carve_size((size_t)n * 4);
  size_t t_words = cac::cc_tmp_words(m, n, label_mode);
  void* base = nullptr;
  rc = cac::scratch_reserve(ctx, e_bytes + l_bytes + cac::carve_size(t_words * 4) + 256, &base);
  if (rc) return rc;
  cac::Carver cv(base);
  cac_uedge* d_e = cv.take<cac_uedge>(m);
  uint32_t* d_l = cv.take<uint32_t>(n);
  uint32_t* tmp = cv.take<uint32_t>(t_words);
  if (m) CAC_HIP(ctx, hipMemcpyAsync(d_e, edges, m * sizeof(cac_uedge), hipMemcpyHostToDevice, s));
  // ids are validated on the device before any union-find access
  hipLaunchKernelGGL(cac::k_zero_u32, dim3(1), dim3(64), 0, s, tmp, 1);
  if (m)
    hipLaunchKernelGGL(cac::k_check_ids, dim3(cac::grid_for(m, 256, 4096)), dim3(256), 0, s, (const cac_uedge*)d_e, m,
                       n, tmp);
  CAC_LAUNCH_CHECK(ctx, "id check");
  CAC_HIP(ctx, hipMemcpyAsync(ctx->pinned, tmp, 4, hipMemcpyDeviceToHost, s));
  CAC_HIP(ctx, hipStreamSynchronize(s));
  if (*(uint32_t*)ctx->pinned) return cac::fail(ctx, CAC_EINVAL, "edge endpoint >= n");
  rc = cac::cc_launch(ctx, d_e, m, n, label_mode, d_l, tmp, s);
  if (rc) return rc;
  CAC_HIP(ctx, hipMemcpyAsync(ctx->pinned, tmp, 4, hipMemcpyDeviceToHost, s));
  if (n) CAC_HIP(ctx, hipMemcpyAsync(labels_out, d_l, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  CAC_HIP(ctx, hipStreamSynchronize(s));
  *ncomp_out = *(uint32_t*)ctx->pinned;
  return CAC_OK;
}

int cac_cc_csr_dev(cac_ctx* ctx, const uint32_t* d_offsets, const uint32_t* d_cols, uint32_t n, int label_mode,
                   uint32_t* d_labels, uint32_t* ncomp_out, void* stream) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if (!d_offsets || (n && (!d_labels || !d_cols)) || !ncomp_out) return cac::fail(ctx, CAC_EINVAL, "null pointer");
  if (label_mode != CAC_LABEL_MIN_ID && label_mode != CAC_LABEL_RANK)
    return cac::fail(ctx, CAC_EINVAL, "label_mode must be CAC_LABEL_MIN_ID or CAC_LABEL_RANK");
  if (n == 0) {
    *ncomp_out = 0;
    return CAC_OK;
  }
  hipStream_t s = cac::pick_stream(ctx, stream);
  void* base = nullptr;
  rc = cac::scratch_reserve(ctx, cac::cc_csr_tmp_words(n, label_mode) * 4, &base);
  if (rc) return rc;
  uint32_t* counter = (uint32_t*)base;
  rc = cac::cc_csr_launch(ctx, d_offsets, d_cols, n, label_mode, d_labels, counter, s);
  if (rc) return rc;
  CAC_HIP(ctx, hipMemcpyAsync(ctx->pinned, counter, 4, hipMemcpyDeviceToHost, s));
  CAC_HIP(ctx, hipStreamSynchronize(s));
  *ncomp_out = *(uint32_t*)ctx->pinned;
  return CAC_OK;
}

int cac_cc_csr_graph_dev(cac_ctx* ctx, const uint32_t* d_offsets, const uint32_t* d_cols, uint32_t n, int label_mode,
                         uint32_t* d_labels, uint32_t* ncomp_out, void* stream) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if (!d_offsets || (n && (!d_labels || !d_cols)) || !ncomp_out) return cac::fail(ctx, CAC_EINVAL, "null pointer");
  if (label_mode != CAC_LABEL_MIN_ID && label_mode != CAC_LABEL_RANK)
    return cac::fail(ctx, CAC_EINVAL, "label_mode must be CAC_LABEL_MIN_ID or CAC_LABEL_RANK");
  if (n == 0) {
    *ncomp_out = 0;
    return CAC_OK;
  }
  hipStream_t s = cac::pick_stream(ctx, stream);
  void* base = nullptr;
  rc = cac::scratch_reserve(ctx, cac::cc_csr_tmp_words(n, label_mode) * 4, &base);
  if (rc) return rc;
  uint32_t* counter = (uint32_t*)base;
  const void* ptrs[5] = {d_offsets, d_cols, d_labels, base, (const void*)s};
  const uint64_t args[2] = {n, (uint64_t)label_mode};
  bool same = ctx->cc_graph != nullptr && args[0] == ctx->cc_graph_args[0] && args[1] == ctx->cc_graph_args[1];
  for (int i = 0; same && i < 5; ++i) same = ptrs[i] == ctx->cc_graph_ptrs[i];
  if (!same) {
    // Capture the launch sequence of cac_cc_csr_dev (kernels only: no memset
    // or copy nodes) with per-kernel timing off, then instantiate it.
    if (ctx->cc_graph) {
      CAC_HIP(ctx, hipGraphExecDestroy(ctx->cc_graph));
      ctx->cc_graph = nullptr;
    }
    const bool prof = ctx->profiling;
    ctx->profiling = false;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      ctx->profiling = prof;
      return cac::hip_fail(ctx, e, "hipStreamBeginCapture");
    }
    rc = cac::cc_csr_launch(ctx, d_offsets, d_cols, n, label_mode, d_labels, counter, s);
    e = hipStreamEndCapture(s, &g);
    ctx->profiling = prof;
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return cac::hip_fail(ctx, e, "hipStreamEndCapture");
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return cac::hip_fail(ctx, e, "hipGraphInstantiate");
    ctx->cc_graph = x;
    for (int i = 0; i < 5; ++i) ctx->cc_graph_ptrs[i] = ptrs[i];
    ctx->cc_graph_args[0] = args[0];
    ctx->cc_graph_args[1] = args[1];
  }
  cac::prof_begin(ctx, "cc_csr_graph", s);
  CAC_HIP(ctx, hipGraphLaunch(ctx->cc_graph, s));
  cac::prof_end(ctx, s);
  CAC_HIP(ctx, hipMemcpyAsync(ctx->pinned, counter, 4, hipMemcpyDeviceToHost, s));
  CAC_HIP(ctx, hipStreamSynchronize(s));
  *ncomp_out = *(uint32_t*)ctx->pinned;
  return CAC_OK;
}

int cac_csr_build_dev(cac_ctx* ctx, const cac_uedge* d_edges, uint64_t m, uint32_t n, uint32_t* d_offsets,
                      uint32_t* d_cols, uint64_t* nnz_out, void* stream) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if ((m && !d_edges) || !d_offsets || !nnz_out || (m && !d_cols)) return cac::fail(ctx, CAC_EINVAL, "null pointer");
  if (2 * m >= 0xFFFFFFFFull) return cac::fail(ctx, CAC_EINVAL, "a CSR needs fewer than 2^31 edges");
  hipStream_t s = cac::pick_stream(ctx, stream);
  uint32_t bits = 1;
  while (bits < 32 && (1ull << bits) < n) ++bits;
  const uint64_t nd = 2 * m;
  uint64_t *k0 = nullptr, *k1 = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  uint32_t* d_nnz = nullptr;
  hipError_t e = hipMalloc((void**)&k0, 8 * (nd ? nd : 1));
  if (e == hipSuccess) e = hipMalloc((void**)&k1, 8 * (nd ? nd : 1));
  if (e == hipSuccess) e = hipMalloc((void**)&d_nnz, 64);
  if (e == hipSuccess && m) {  // ids < n, checked before any offset is written
    uint32_t bad = 0;
    hipLaunchKernelGGL(cac::k_zero_u32, dim3(1), dim3(64), 0, s, d_nnz, 1);
    hipLaunchKernelGGL(cac::k_check_ids, dim3(cac::grid_for(m, 256, 4096)), dim3(256), 0, s, d_edges, m, n, d_nnz);
    e = hipMemcpyAsync(&bad, d_nnz, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && bad) {
      (void)hipFree(d_nnz);
      (void)hipFree(k1);
      (void)hipFree(k0);
      return cac::fail(ctx, CAC_EINVAL, "edge endpoint >= n");
    }
  }
  rocprim::double_buffer<uint64_t> kb(k0, k1);
  if (e == hipSuccess && nd) e = rocprim::radix_sort_keys(nullptr, tmp_bytes, kb, nd, 0, 2 * bits, s);
  if (e == hipSuccess && nd) e = hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1);
  uint64_t nnz = 0;
  if (e == hipSuccess && nd) {
    hipLaunchKernelGGL(cac::k_csr_keys, dim3(cac::grid_for(m, 256, 8192)), dim3(256), 0, s, d_edges, m, bits, k0);
    e = rocprim::radix_sort_keys(tmp, tmp_bytes, kb, nd, 0, 2 * bits, s);
  }
  if (e == hipSuccess && nd) {
    // loops sort last (all ones): nnz = number of keys below the loop marker,
    // by binary search over single-key reads
    uint64_t lo = 0, hi = nd;
    while (lo < hi && e == hipSuccess) {
      const uint64_t mid = lo + (hi - lo) / 2;
      uint64_t key = 0;
      e = hipMemcpyAsync(&key, kb.current() + mid, 8, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (key == ~0ull) hi = mid;
      else lo = mid + 1;
    }
    nnz = lo;
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(cac::k_csr_split, dim3(cac::grid_for(nnz + 1, 256, 8192)), dim3(256), 0, s,
                       (const uint64_t*)kb.current(), nnz, n, bits, d_offsets, d_cols);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) rc = cac::hip_fail(ctx, e, "csr build");
  if (tmp) (void)hipFree(tmp);
  if (d_nnz) (void)hipFree(d_nnz);
  if (k1) (void)hipFree(k1);
  if (k0) (void)hipFree(k0);
  if (rc == CAC_OK) *nnz_out = nnz;
  return rc;
}

}  // extern "C"

namespace cac {
int approx_cut_run(cac_ctx* ctx, const cac_edge* d_edges, uint64_t m, uint32_t n, uint32_t copies, uint32_t iters,
                   uint32_t seed, uint32_t* first_disconnected, hipStream_t s) {
  const uint64_t N = (uint64_t)n * copies;
  if (N >= 0xFFFFFFFFull) return fail(ctx, CAC_EINVAL, "n * trials must fit in 32-bit vertex ids");
  void* base = nullptr;
  int rc = scratch_reserve(ctx, carve_size(N * 4) + 512, &base);
  if (rc) return rc;
  Carver cv(base);
  uint32_t* par = cv.take<uint32_t>(N);
  uint32_t* counter = cv.take<uint32_t>(64);
  *first_disconnected = 0;
  for (uint32_t i = 1; i <= iters; ++i) {
    hipLaunchKernelGGL(k_cc_init, dim3(grid_for(N, kCcBlock, 256 * 64)), dim3(kCcBlock), 0, s, par, (uint32_t)N);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, counter, 1);
    prof_begin(ctx, "approx_hook", s);
    hipLaunchKernelGGL(k_approx_hook, dim3(grid_for(m, kCcBlock, 256 * 32)), dim3(kCcBlock), 0, s, d_edges, m, n,
                       copies, i, seed, par);
    prof_end(ctx, s);
    hipLaunchKernelGGL(k_cc_compress, dim3(grid_for(N, kCcBlock, 4096)), dim3(kCcBlock), 0, s, par, (uint32_t)N,
                       counter);
    CAC_LAUNCH_CHECK(ctx, "approx kernels");
    CAC_HIP(ctx, hipMemcpyAsync(ctx->pinned, counter, 4, hipMemcpyDeviceToHost, s));
    CAC_HIP(ctx, hipStreamSynchronize(s));
    if (*(uint32_t*)ctx->pinned > copies) {
      *first_disconnected = i;
      break;
    }
  }
  return CAC_OK;
}
}  // namespace cac
