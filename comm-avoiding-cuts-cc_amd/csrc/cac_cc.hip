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
// larger root to point at the smaller.
__device__ inline void uf_link_from(uint32_t* par, uint32_t a, uint32_t pa, uint32_t b, uint32_t pb) {
  uint32_t ra = uf_find_from(par, a, pa), rb = uf_find_from(par, b, pb);
  while (ra != rb) {
    uint32_t lo = ra < rb ? ra : rb;
    uint32_t hi = ra < rb ? rb : ra;
    uint32_t old = atomicCAS(par + hi, hi, lo);
    if (old == hi) return;
    // hi was hooked elsewhere meanwhile: climb from its new parent and retry.
    ra = uf_find(par, old);
    rb = uf_find(par, lo);
  }
}

__device__ inline void uf_link(uint32_t* par, uint32_t a, uint32_t b) {
  uf_link_from(par, a, ld_par(par, a), b, ld_par(par, b));
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

__global__ void __launch_bounds__(kCcBlock) k_cc_settle(uint32_t* par, uint32_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * kCcBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kCcBlock) {
    uint32_t v = (uint32_t)i;
    uint32_t p = ld_par_fresh(par, v);
    if (p == v) continue;
    uint32_t r = uf_find_fresh(par, p);
    if (r != p) st_par(par, v, r);
  }
}

// Most frequent root among kModeSamples vertices spread over [0, n) (parents
// are settled, so par[v] is v's root).  Only a performance choice: any root
// gives correct labels.
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
    const uint32_t v = (uint32_t)(((uint64_t)h * n) >> 32);
    const uint32_t r = ld_par_fresh(par, v);
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
  uint32_t roots = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * kCcBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kCcBlock) {
    uint32_t v = (uint32_t)i;
    uint32_t p = ld_par_fresh(par, v);
    if (p == v) {
      ++roots;
      continue;
    }
    uint32_t r = uf_find_fresh(par, p);
    if (r != p) st_par(par, v, r);
  }
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
