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
//   2. sample    hook one 256-edge chunk out of every kSampleStride chunks; this
//                builds the giant component from ~1/16 of the edges
//   3. settle    pointer-jump every vertex to its root
//   4. hook      stream every remaining edge (8 bytes each, non-temporal) and
//                skip it when both endpoints already share a parent; otherwise
//                union-find link
//   5. compress  parent[v] = root(v), count roots               (8n bytes)
//   6. (rank)    dense rank of roots by id, relabel              (scan)
#include "cac_common.h"
#include "cac_internal.h"
#include "cac_scan.h"

namespace cac {
namespace {

constexpr int kCcBlock = 256;
constexpr int kChunk = 256;          // edges per sampling chunk (one per thread of a block)
constexpr int kSampleStride = 16;    // 1 sampled chunk per kSampleStride chunks
constexpr int kEdgesPerThread = 8;   // edges each hook thread keeps in flight

// Parent-array accesses are relaxed atomics at workgroup scope: plain cached
// loads/stores (agent scope would add sc1 and bypass the per-XCD L2 on every
// hop).  A stale parent read from another XCD's older state is still an
// ancestor, so finds stay correct; only the CAS in uf_link needs agent-wide
// coherence, and kernel boundaries publish every write.
__device__ inline uint32_t ld_par(const uint32_t* p, uint32_t i) {
  return __hip_atomic_load(p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void st_par(uint32_t* p, uint32_t i, uint32_t v) {
  __hip_atomic_store(p + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
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

// One block per sampled chunk: chunk c = blockIdx.x * kSampleStride.
__global__ void __launch_bounds__(kCcBlock) k_cc_sample(const cac_uedge* __restrict__ edges, uint64_t m,
                                                        uint32_t* par) {
  uint64_t i = ((uint64_t)blockIdx.x * kSampleStride) * kChunk + threadIdx.x;
  if (i >= m) return;
  uint64_t e = ld_edge_nt(edges, i);
  uint32_t u = (uint32_t)e, v = (uint32_t)(e >> 32);
  if (u != v) uf_link(par, u, v);
}

__global__ void __launch_bounds__(kCcBlock) k_cc_settle(uint32_t* par, uint32_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * kCcBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kCcBlock) {
    uint32_t v = (uint32_t)i;
    uint32_t p = ld_par(par, v);
    if (p == v) continue;
    uint32_t r = uf_find(par, v);
    if (r != p) st_par(par, v, r);
  }
}

// Every chunk not hooked by k_cc_sample.  Thread t of the grid owns edges
// t, t + T, t + 2T, ... (T = total threads) so each load instruction is
// coalesced; kEdgesPerThread loads are issued before any is consumed.
__global__ void __launch_bounds__(kCcBlock) k_cc_hook(const cac_uedge* __restrict__ edges, uint64_t m,
                                                      uint32_t* par, int skip_sampled) {
  const uint64_t T = (uint64_t)gridDim.x * kCcBlock;
  uint64_t i0 = (uint64_t)blockIdx.x * kCcBlock + threadIdx.x;
  for (; i0 < m; i0 += T * kEdgesPerThread) {
    uint64_t e[kEdgesPerThread];
#pragma unroll
    for (int k = 0; k < kEdgesPerThread; ++k) {
      uint64_t i = i0 + (uint64_t)k * T;
      bool sampled = skip_sampled && ((i / kChunk) % kSampleStride) == 0;
      e[k] = (i < m && !sampled) ? ld_edge_nt(edges, i) : 0ull;
    }
    // all parent lookups of the batch are in flight before any is consumed
    uint32_t pu[kEdgesPerThread], pv[kEdgesPerThread];
#pragma unroll
    for (int k = 0; k < kEdgesPerThread; ++k) {
      pu[k] = ld_par(par, (uint32_t)e[k]);
      pv[k] = ld_par(par, (uint32_t)(e[k] >> 32));
    }
#pragma unroll
    for (int k = 0; k < kEdgesPerThread; ++k) {
      if (pu[k] == pv[k]) continue;  // also covers loops and empty slots (u == v)
      uf_link_from(par, (uint32_t)e[k], pu[k], (uint32_t)(e[k] >> 32), pv[k]);
    }
  }
}

// Final pointer jumping; counts roots (components).
__global__ void __launch_bounds__(kCcBlock) k_cc_compress(uint32_t* par, uint32_t n, uint32_t* ncomp) {
  uint32_t roots = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * kCcBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kCcBlock) {
    uint32_t v = (uint32_t)i;
    uint32_t p = ld_par(par, v);
    if (p == v) {
      ++roots;
      continue;
    }
    uint32_t r = p, q = ld_par(par, r);
    while (q < r) {
      r = q;
      q = ld_par(par, r);
    }
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

// Launch sequence shared by cac_cc and cac_cc_dev.  'counter' is device memory
// for one u32; 'rank_tmp' holds n + scan partials when label_mode == RANK.
int cc_launch(cac_ctx* ctx, const cac_uedge* d_edges, uint64_t m, uint32_t n, int label_mode, uint32_t* d_lab,
              uint32_t* d_counter, uint32_t* d_rank_tmp, hipStream_t s) {
  const unsigned cap = 256 * 64;
  prof_begin(ctx, "cc_init", s);
  hipLaunchKernelGGL(k_cc_init, dim3(grid_for(n, kCcBlock, cap)), dim3(kCcBlock), 0, s, d_lab, n);
  prof_end(ctx, s);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, d_counter, 1);
  const uint64_t chunks = (m + kChunk - 1) / kChunk;
  const uint64_t sampled = (chunks + kSampleStride - 1) / kSampleStride;
  const bool use_sample = m >= (uint64_t)kChunk * kSampleStride * 64;
  if (use_sample) {
    prof_begin(ctx, "cc_sample", s);
    hipLaunchKernelGGL(k_cc_sample, dim3((unsigned)sampled), dim3(kCcBlock), 0, s, d_edges, m, d_lab);
    prof_end(ctx, s);
    prof_begin(ctx, "cc_settle", s);
    hipLaunchKernelGGL(k_cc_settle, dim3(grid_for(n, kCcBlock, cap)), dim3(kCcBlock), 0, s, d_lab, n);
    prof_end(ctx, s);
  }
  if (m) {
    prof_begin(ctx, "cc_hook", s);
    unsigned g = grid_for((m + kEdgesPerThread - 1) / kEdgesPerThread, kCcBlock, 256 * 32);
    hipLaunchKernelGGL(k_cc_hook, dim3(g), dim3(kCcBlock), 0, s, d_edges, m, d_lab, use_sample ? 1 : 0);
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

size_t cc_rank_tmp_words(uint32_t n) { return ((size_t)(n + 63) & ~size_t(63)) + scan_tiles(n) + 64; }

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
  size_t words = 64 + (label_mode == CAC_LABEL_RANK ? cac::cc_rank_tmp_words(n) : 0);
  void* base = nullptr;
  rc = cac::scratch_reserve(ctx, words * 4, &base);
  if (rc) return rc;
  uint32_t* counter = (uint32_t*)base;
  rc = cac::cc_launch(ctx, d_edges, m, n, label_mode, d_labels, counter, counter + 64, s);
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
  size_t t_words = 64 + (label_mode == CAC_LABEL_RANK ? cac::cc_rank_tmp_words(n) : 0);
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
  rc = cac::cc_launch(ctx, d_e, m, n, label_mode, d_l, tmp, tmp + 64, s);
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
