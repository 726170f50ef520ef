// cac_sample.hip -- weighted iterated sampling (the sqrt-cut trial's shrink
// phase) and the low-concurrency sqrt-cut driver on gfx950.
//
// Replaces SequentialSquareRootCut::iteratedSampling / compute
// (src/SequentialSquareRootCut.cpp:4-67) with AdjacencyListGraph::
// {weaklyContractEdge, finalizeContractionPhase, compact}
// (src/AdjacencyListGraph.cpp:34-112) and the trial loop of
// SquareRootCut::runLowConcurrencyMaster (src/SquareRootCut.cpp:86-148).
//
// One workgroup runs one trial's rounds (trials run side by side):
//   1. prefix of the edge weights with loops masked to 0 (block scan; u64,
//      or u32 once a previous round's total fits in 32 bits);
//   2. |V| draws: Philox(ctr = {trial, round, draw, 0}, key = {seed, 'SAMP'}),
//      r = 1 + floor(x W / 2^64), edge = first prefix >= r (binary search);
//   3. "contract in draw order until |V| == t" (a prefix condition): binary
//      search over the prefix length with lock-free union-find probes, each
//      counting its successful hooks;
//   4. canonical relabel: every vertex takes the smallest original id of its
//      super-vertex (hooks always attach the larger root under the smaller).
// Densification then ranks the canonical labels and accumulates the weights of
// the surviving edges into a t x t int64 matrix (the merge of parallel edges),
// which feeds one dense Karger-Stein trial per sqrt-cut trial.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <unordered_map>
#include <vector>

#include <cstring>

#include <rocprim/rocprim.hpp>

#include "cac_common.h"
#include "cac_internal.h"
#include "cac_ks.h"
#include "cac_ks_kernels.h"
#include "cac_scan.h"

namespace cac {
namespace {

constexpr int kSmpBlock = 1024;
constexpr int kSmpWaves = kSmpBlock / 64;

// One sampling job = one sqrt-cut trial of one graph; a launch runs one
// workgroup per job, so trials of different graphs share a launch.
struct TrialScratch {
  const cac_edge* edges;  // the job's graph (ids local to it)
  uint64_t m;
  uint32_t n, t;          // vertices, intermediate size
  uint32_t trial, slot;   // Philox trial id, result slot
  uint32_t* lab;          // [n]
  uint32_t* par;          // [n]
  uint64_t* pre;          // [m]
  uint2* draws;           // [n]
  int64_t* mat;           // [t x t]
  int64_t* rsum;          // [t] row sums of mat (may be null)
  uint32_t* nv_out;       // [1]
  // [1] of the job's graph, nonzero when an edge is a loop or two weights
  // differ (may be null: then every round builds the prefix)
  const uint32_t* mixed;
};

// A distinct graph of a sampling launch and its flag word.
struct EdgeSpan {
  const cac_edge* edges;
  uint64_t m;
  uint32_t* mixed;
};

__global__ void k_mixed_zero(const EdgeSpan* spans, uint32_t count) {
  for (uint32_t g = blockIdx.x * 256 + threadIdx.x; g < count; g += gridDim.x * 256)
    __hip_atomic_store(spans[g].mixed, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid (slices, graphs): flags graph blockIdx.y when one of its edges is a loop
// or weighs other than its first edge (or the first weighs 0).
__global__ void __launch_bounds__(256) k_mixed_weights(const EdgeSpan* spans) {
  const EdgeSpan sp = spans[blockIdx.y];
  if (sp.m == 0) return;
  const uint64_t w0 = sp.edges[0].w;
  bool bad = w0 == 0;
  for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < sp.m; e += (uint64_t)gridDim.x * 256) {
    const cac_edge x = sp.edges[e];
    bad = bad || x.w != w0 || x.u == x.v;
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(sp.mixed, 1u);
}

__device__ inline uint64_t wsum64(uint64_t x) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

__device__ inline uint64_t bsum64(uint64_t x, uint64_t* lds) {
  x = wsum64(x);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = x;
  __syncthreads();
  uint64_t t = 0;
  for (int w = 0; w < kSmpWaves; ++w) t += lds[w];
  __syncthreads();
  return t;
}

__device__ inline uint32_t uf_root(uint32_t* par, uint32_t x) {
  uint32_t p = par[x];
  while (p != x) {
    uint32_t g = par[p];
    if (g != p) par[x] = g;
    x = p;
    p = g;
  }
  return x;
}

// Hook the sets of a and b; true when this call merged two roots.
__device__ inline bool uf_union(uint32_t* par, uint32_t a, uint32_t b) {
  uint32_t ra = uf_root(par, a), rb = uf_root(par, b);
  while (ra != rb) {
    uint32_t lo = min(ra, rb), hi = max(ra, rb);
    uint32_t old = atomicCAS(par + hi, hi, lo);
    if (old == hi) return true;
    ra = uf_root(par, old);
    rb = uf_root(par, lo);
  }
  return false;
}

// Union-find over the first L draws; returns the hooks made.  Every label of
// the round starts as its own root: at the round's first probe they are the
// roots the previous round left (undo = 0), and a later probe first undoes the
// previous one, whose first `undo` draws' endpoints are the only entries it
// changed (hooks join roots found from endpoints, and path splitting only
// re-points nodes on their paths, all endpoints).
__device__ uint32_t probe(const uint2* draws, uint32_t L, uint32_t undo, uint32_t* par, uint64_t* lds) {
  for (uint32_t d = threadIdx.x; d < undo; d += kSmpBlock) {
    const uint2 x = draws[d];
    par[x.x] = x.x;
    par[x.y] = x.y;
  }
  block_sync_global();  // no wave may see the previous probe's forest
  uint64_t mine = 0;
  for (uint32_t d = threadIdx.x; d < L; d += kSmpBlock) mine += uf_union(par, draws[d].x, draws[d].y) ? 1 : 0;
  block_sync_global();  // the relabel reads roots after the last hook
  return (uint32_t)bsum64(mine, lds);
}

// kMinWaves = 8 holds the kernel to 64 VGPRs (8 waves per SIMD), so two 16-wave
// workgroups share a CU, at the price of a few spilled registers: for launches
// of more workgroups than CUs.  kMinWaves = 1 leaves the registers free, for
// small launches whose trials each have a CU to themselves.
template <int kMinWaves>
__global__ void __launch_bounds__(kSmpBlock, kMinWaves)
    k_sample_trials(uint32_t seed, const TrialScratch* __restrict__ scratch) {
  __shared__ uint64_t lds[kSmpWaves + 2];
  const TrialScratch S = scratch[blockIdx.x];
  const cac_edge* __restrict__ edges = S.edges;
  const uint64_t m = S.m;
  const uint32_t n = S.n, t = S.t, trial = S.trial;
  for (uint32_t v = threadIdx.x; v < n; v += kSmpBlock) {
    S.lab[v] = v;
    S.par[v] = v;
  }
  block_sync_global();
  // One weight wu on every edge and no loop: the first round's prefix is
  // wu * (e + 1), so its draws need neither the prefix nor a search.
  uint64_t wu = 0;
  if (S.mixed && m > 0 && __hip_atomic_load(S.mixed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
    wu = edges[0].w;
  uint32_t nv = n;
  // A round's masked total never exceeds the previous round's (edges only turn
  // into loops), so once a total fits in 32 bits the prefix is kept as u32:
  // half the prefix bytes, the same comparisons.
  uint64_t prev_w = ~0ull;
  uint32_t* pre32 = reinterpret_cast<uint32_t*>(S.pre);
  for (uint32_t round = 0; nv > t; ++round) {
    const bool flat = round == 0 && wu != 0;
    const bool narrow = prev_w <= 0xffffffffull;
    // 1. masked inclusive prefix of the weights
    uint64_t carry = flat ? wu * m : 0;
    constexpr int kItems = 4;
    for (uint64_t base = 0; !flat && base < m; base += (uint64_t)kSmpBlock * kItems) {
      const uint64_t e0 = base + (uint64_t)threadIdx.x * kItems;
      uint64_t w[kItems], s = 0;
#pragma unroll
      for (int k = 0; k < kItems; ++k) {
        const uint64_t e = e0 + k;
        w[k] = 0;
        if (e < m) {
          const cac_edge x = edges[e];
          if (S.lab[x.u] != S.lab[x.v]) w[k] = x.w;
        }
        s += w[k];
      }
      // block exclusive scan of s
      const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
      uint64_t inc = s;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        uint64_t y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
      }
      __shared__ uint64_t wtot[kSmpWaves];
      if (lane == 63) wtot[wave] = inc;
      __syncthreads();
      uint64_t off = carry, tot = 0;
      for (int v = 0; v < kSmpWaves; ++v) {
        const uint64_t q = wtot[v];
        if (v < wave) off += q;
        tot += q;
      }
      uint64_t run = off + inc - s;
#pragma unroll
      for (int k = 0; k < kItems; ++k) {
        const uint64_t e = e0 + k;
        run += w[k];
        if (e < m) {
          if (narrow) pre32[e] = (uint32_t)run;
          else S.pre[e] = run;
        }
      }
      carry += tot;
      __syncthreads();
    }
    block_sync_global();  // the draws read the prefix written by every wave
    const uint64_t W = carry;
    prev_w = W;
    if (W == 0) break;  // only loops left: more than t components
    // 2. draws
    const uint32_t D = nv;
    for (uint32_t d = threadIdx.x; d < D; d += kSmpBlock) {
      const u32x4 rnd = philox4x32(trial, round, d, 0u, seed, kDomSample);
      const uint64_t r = uniform1(lo64(rnd), W);
      uint64_t lo = flat ? (r - 1) / wu : 0, hi = flat ? lo : m - 1;
      while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if ((narrow ? (uint64_t)pre32[mid] : S.pre[mid]) >= r) hi = mid;
        else lo = mid + 1;
      }
      const cac_edge x = edges[lo];
      S.draws[d] = make_uint2(S.lab[x.u], S.lab[x.v]);
    }
    block_sync_global();
    // 3. shortest draw prefix that leaves t super-vertices
    uint32_t hooks = probe(S.draws, D, 0, S.par, lds);
    if (nv - hooks <= t) {
      uint32_t lo = 1, hi = D, last = D;
      while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        hooks = probe(S.draws, mid, last, S.par, lds);
        last = mid;
        if (nv - hooks <= t) hi = mid;
        else lo = mid + 1;
      }
      if (last != lo) hooks = probe(S.draws, lo, last, S.par, lds);
    }  // else all D draws are kept, and the forest of the first probe stays
    // 4. canonical relabel
    for (uint32_t v = threadIdx.x; v < n; v += kSmpBlock) S.lab[v] = uf_root(S.par, S.lab[v]);
    block_sync_global();
    nv -= hooks;
  }
  if (threadIdx.x == 0) *S.nv_out = nv;
}

// Dynamic LDS of k_densify_trials: the row sums of a matrix of up to kKsMaxS rows.
constexpr size_t kDensifyLds = (size_t)kKsMaxS * 8;
constexpr int kDensifyEdges = 4;  // edges per thread in flight

// Rank the canonical labels, accumulate the surviving edges into the k x k
// matrix (k == t unless the sampling stopped early on a disconnected graph),
// with its row sums in LDS, and compute the trial's bound (minimum weighted
// degree).  bounds is indexed by job (blockIdx.x), results by the job's slot.
__global__ void __launch_bounds__(kSmpBlock) k_densify_trials(const TrialScratch* __restrict__ scratch, int64_t* bounds,
                                                              int64_t* results) {
  __shared__ uint32_t wtot[kSmpWaves];
  __shared__ uint64_t lds[kSmpWaves + 2];
  extern __shared__ unsigned long long row_sum[];  // [k]
  const TrialScratch S = scratch[blockIdx.x];
  const cac_edge* __restrict__ edges = S.edges;
  const uint64_t m = S.m;
  const uint32_t n = S.n, t = S.t, slot = S.slot;
  const uint32_t k = *S.nv_out;
  if (k > t) {  // disconnected: a cut of weight 0 exists
    if (threadIdx.x == 0) {
      bounds[blockIdx.x] = 0;
      results[slot] = 0;
    }
    return;
  }
  uint32_t* rank = S.par;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < n; base += kSmpBlock) {
    const uint32_t x = base + threadIdx.x;
    const uint32_t live = (x < n && S.lab[x] == x) ? 1u : 0u;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = live;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      uint32_t y = __shfl_up(inc, d, 64);
      if (lane >= d) inc += y;
    }
    if (lane == 63) wtot[wave] = inc;
    __syncthreads();
    uint32_t off = carry, tot = 0;
    for (int v = 0; v < kSmpWaves; ++v) {
      if (v < wave) off += wtot[v];
      tot += wtot[v];
    }
    if (live) rank[x] = off + inc - 1;
    carry += tot;
    __syncthreads();
  }
  for (uint64_t i = threadIdx.x; i < (uint64_t)k * k; i += kSmpBlock) S.mat[i] = 0;
  const bool lds_rows = k <= kKsMaxS;  // else the row sums are read back from the matrix
  for (uint32_t r = threadIdx.x; lds_rows && r < k; r += kSmpBlock) row_sum[r] = 0;
  block_sync_global();  // ranks and zeroed rows before the accumulation
  // kDensifyEdges edges per thread: their label and rank loads are in flight together
  for (uint64_t base = 0; base < m; base += (uint64_t)kSmpBlock * kDensifyEdges) {
    cac_edge x[kDensifyEdges];
    uint32_t a[kDensifyEdges], b[kDensifyEdges];
#pragma unroll
    for (int j = 0; j < kDensifyEdges; ++j) {
      const uint64_t e = base + (uint64_t)j * kSmpBlock + threadIdx.x;
      x[j] = e < m ? edges[e] : cac_edge{0u, 0u, 0ull};  // vertex 0 stands in: same label, skipped
    }
#pragma unroll
    for (int j = 0; j < kDensifyEdges; ++j) {
      a[j] = S.lab[x[j].u];
      b[j] = S.lab[x[j].v];
    }
#pragma unroll
    for (int j = 0; j < kDensifyEdges; ++j) {
      a[j] = rank[a[j]];
      b[j] = rank[b[j]];
    }
#pragma unroll
    for (int j = 0; j < kDensifyEdges; ++j) {
      if (a[j] == b[j]) continue;
      const unsigned long long w = x[j].w;
      atomicAdd((unsigned long long*)&S.mat[(uint64_t)a[j] * k + b[j]], w);
      atomicAdd((unsigned long long*)&S.mat[(uint64_t)b[j] * k + a[j]], w);
      if (lds_rows) {
        atomicAdd(&row_sum[a[j]], w);
        atomicAdd(&row_sum[b[j]], w);
      }
    }
  }
  block_sync_global();
  // row sums (the Karger-Stein capacities) and bound = min row sum
  uint64_t best = ~0ull;
  if (lds_rows) {
    for (uint32_t r = threadIdx.x; r < k; r += kSmpBlock) {
      const uint64_t acc = row_sum[r];
      if (S.rsum) S.rsum[r] = (int64_t)acc;
      best = acc < best ? acc : best;
    }
  } else {  // one wave per row
    for (uint32_t r = threadIdx.x >> 6; r < k; r += kSmpWaves) {
      uint64_t acc = 0;
      for (uint32_t x = threadIdx.x & 63; x < k; x += 64) acc += (uint64_t)S.mat[(uint64_t)r * k + x];
      acc = wsum64(acc);
      if (S.rsum && (threadIdx.x & 63) == 0) S.rsum[r] = (int64_t)acc;
      best = acc < best ? acc : best;
    }
  }
  for (int d = 32; d > 0; d >>= 1) {
    const uint64_t o = __shfl_xor(best, d, 64);
    best = o < best ? o : best;
  }
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t b = ~0ull;
    for (int w = 0; w < kSmpWaves; ++w) b = lds[w] < b ? lds[w] : b;
    bounds[blockIdx.x] = (int64_t)b;
    results[slot] = (int64_t)b;
  }
}

// ---- trials whose contracted graph exceeds the dense engine (t > 8192) ------
// The contracted graph becomes an edge list, as SequentialKargerSteinCut gets
// it (AdjacencyListGraph::compact, src/AdjacencyListGraph.cpp:34-112): endpoints
// ranked by canonical label, loops dropped, parallel edges summed, sorted by
// (u, v); one sparse Karger-Stein trial runs on it.
struct LabLive {
  const uint32_t* lab;
  __device__ uint32_t operator()(uint64_t i) const { return lab[i] == (uint32_t)i ? 1u : 0u; }
};
struct LabRank {
  uint32_t* rank;
  __device__ void operator()(uint64_t i, uint32_t ex, uint32_t live) const {
    if (live) rank[i] = ex;
  }
};

// key (a << 32) | b of the ranked endpoints (a < b), ~0 for a loop; value w.
__global__ void __launch_bounds__(256) k_lt_keys(const cac_edge* __restrict__ edges, uint64_t m,
                                                 const uint32_t* __restrict__ lab, const uint32_t* __restrict__ rank,
                                                 uint64_t* __restrict__ keys, uint64_t* __restrict__ vals) {
  for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < m; e += (uint64_t)gridDim.x * 256) {
    const cac_edge x = edges[e];
    const uint32_t a = rank[lab[x.u]], b = rank[lab[x.v]];
    keys[e] = a == b ? ~0ull : ((uint64_t)min(a, b) << 32) | max(a, b);
    vals[e] = x.w;
  }
}

// Merged keys to edges, and the weighted degrees.
__global__ void __launch_bounds__(256) k_lt_edges(const uint64_t* __restrict__ ukeys, const uint64_t* __restrict__ sums,
                                                  uint64_t count, cac_edge* __restrict__ out,
                                                  unsigned long long* __restrict__ deg) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (uint64_t)gridDim.x * 256) {
    const uint64_t k = ukeys[i];
    const uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)k;
    out[i] = cac_edge{a, b, sums[i]};
    atomicAdd(&deg[a], (unsigned long long)sums[i]);
    atomicAdd(&deg[b], (unsigned long long)sums[i]);
  }
}

__global__ void __launch_bounds__(256) k_lt_zero(unsigned long long* p, uint64_t count) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (uint64_t)gridDim.x * 256) p[i] = 0;
}

// *out = min(deg[0..k)) (out preset to ~0), and results[slot] = that bound.
__global__ void __launch_bounds__(256) k_lt_min(const unsigned long long* deg, uint32_t k, unsigned long long* out) {
  __shared__ unsigned long long w[4];
  unsigned long long b = ~0ull;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < k; i += (uint64_t)gridDim.x * 256) b = min(b, deg[i]);
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) b = min(b, (unsigned long long)__shfl_xor(b, d, 64));
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = b;
  __syncthreads();
  if (threadIdx.x == 0) atomicMin(out, min(min(w[0], w[1]), min(w[2], w[3])));
}

__global__ void k_lt_set(int64_t* p, const unsigned long long* v) {
  if (threadIdx.x == 0) p[0] = (int64_t)v[0];
}


// Points scr[i].mixed at a flag word of its graph (one per distinct edge list)
// and launches the kernels that set the flags; spans and flags hold scr.size()
// entries.  The flags are ready for any later launch on s.
int launch_mixed_flags(cac_ctx* ctx, std::vector<TrialScratch>& scr, EdgeSpan* d_spans, uint32_t* d_flags,
                       hipStream_t s) {
  std::unordered_map<const cac_edge*, uint32_t> index;
  std::vector<EdgeSpan> spans;
  uint64_t mmax = 0;
  for (TrialScratch& x : scr) {
    auto it = index.find(x.edges);
    uint32_t g = 0;
    if (it == index.end()) {
      g = (uint32_t)spans.size();
      index.emplace(x.edges, g);
      spans.push_back({x.edges, x.m, d_flags + g});
      mmax = std::max(mmax, x.m);
    } else {
      g = it->second;
    }
    x.mixed = d_flags + g;
  }
  if (spans.empty()) return CAC_OK;
  const uint32_t count = (uint32_t)spans.size();
  CAC_HIP(ctx, hipMemcpyAsync(d_spans, spans.data(), count * sizeof(EdgeSpan), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_mixed_zero, dim3(grid_for(count, 256, 64)), dim3(256), 0, s, (const EdgeSpan*)d_spans, count);
  for (uint32_t g0 = 0; g0 < count; g0 += 65535u)  // grid y holds at most 65535 graphs
    hipLaunchKernelGGL(k_mixed_weights, dim3(grid_for(mmax, 256 * 16, 256), std::min(count - g0, 65535u)), dim3(256),
                       0, s, (const EdgeSpan*)(d_spans + g0));
  CAC_LAUNCH_CHECK(ctx, "weight flags");
  return CAC_OK;
}


// The densify kernel's dynamic LDS exceeds the default limit: raise it once.
int densify_attr(cac_ctx* ctx) {
  static std::once_flag once;
  static hipError_t err = hipSuccess;
  std::call_once(once, [] {
    err = hipFuncSetAttribute((const void*)k_densify_trials, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kDensifyLds);
  });
  CAC_HIP(ctx, err);
  return CAC_OK;
}


// Sampling workgroups the device runs at once (workgroups per CU times CUs).
size_t sampling_wave(cac_ctx* ctx) {
  if (ctx->smp_wave == 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_sample_trials<8>, kSmpBlock, 0) !=
            hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess) {
      (void)hipGetLastError();
      per_cu = 1;
      cus = 1;
    }
    ctx->smp_wave = (size_t)std::max(per_cu, 1) * (size_t)std::max(cus, 1);
  }
  return ctx->smp_wave;
}


// One sampling launch of nb trials, in the form suited to its size.
void launch_sampling(cac_ctx* ctx, size_t nb, uint32_t seed, const TrialScratch* d_scr, hipStream_t s) {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess) {
    (void)hipGetLastError();
    cus = 1;
  }
  if (nb > (size_t)cus)
    hipLaunchKernelGGL(k_sample_trials<8>, dim3((unsigned)nb), dim3(kSmpBlock), 0, s, seed, d_scr);
  else
    hipLaunchKernelGGL(k_sample_trials<1>, dim3((unsigned)nb), dim3(kSmpBlock), 0, s, seed, d_scr);
}

}  // namespace

// Host rules, same floating point as the reference (SquareRootCut.cpp:63-83).
uint32_t sqrt_intermediate_size(uint32_t n, uint64_t m) {
  uint32_t c = (uint32_t)std::ceil(2.0 * std::sqrt((double)(uint32_t)m) + 1);
  return c > n ? n : c;
}

static double ks_min_success_host(uint32_t s) {
  if (s <= kKsBase) return 1.0;
  double r = ks_min_success_host(ks_target_host(s));
  return 1.0 - std::pow((1 - 0.5 * r), 2);
}

uint32_t ks_trials_host(uint32_t n, double p) {
  if (n <= kKsBase || p == 0) return 1;
  double s = ks_min_success_host(n), cur = 1.0, target = 1.0 - p;
  uint32_t k = 0;
  while (cur > target) {
    cur = cur * (1.0 - s);
    ++k;
  }
  return k;
}

uint32_t sqrt_trials_host(uint32_t n, uint64_t m, double p) {
  double pk = ks_min_success_host(sqrt_intermediate_size(n, m));
  double cp = 1.0 / (1 - p);
  return (uint32_t)std::ceil((n * (n * std::log(cp))) / (2.0 * 2.0 * ((uint32_t)m * pk)));
}

// Device bytes one sampling job needs beside the Karger-Stein levels.
size_t job_sample_bytes(const SqrtJob& j) {
  return carve_size((size_t)j.n * 4) * 2 + carve_size(std::max<uint64_t>(j.m, 1) * 8) + carve_size((size_t)j.n * 8) +
         carve_size((size_t)j.t * j.t * 8) + carve_size((size_t)j.t * 8) + 256;
}
// Upper estimate of the Karger-Stein level arenas of one root of size t: each
// level holds about t^2 entries in copy-on-write work rows plus the contracted
// children, and two levels are alive at once.
size_t job_ks_bytes(const SqrtJob& j) { return 2 * ks_level_bytes_per_root(j.t); }

// Runs sqrt-cut jobs (graph, trial) in chunks sized to the free device memory:
// one sampling launch, one densify launch and one breadth-first Karger-Stein
// pass per chunk, whatever graphs the chunk's jobs belong to.  Raw cut of job
// i goes to d_results[jobs[i].slot].
// Trials with t > 8192, in chunks sized to the free memory: one sampling launch
// for the chunk, then per trial the merged contracted edge list (rank, keys,
// radix sort, reduce by key; the sort buffers are shared) and its bound, then
// one sparse Karger-Stein pass over all of the chunk's trials (node 1 each).
// Raw cut of each job to d_results[slot].
static int sqrt_large_batch(cac_ctx* ctx, const std::vector<SqrtJob>& jobs, uint32_t seed, int64_t* d_results,
                            hipStream_t s) {
  auto job_bytes = [](const SqrtJob& j) {
    const uint64_t m = std::max<uint64_t>(j.m, 1);
    return carve_size(sizeof(TrialScratch)) + 3 * carve_size((size_t)j.n * 4) + carve_size(m * 8) +
           carve_size((size_t)j.n * 8) + carve_size(64) + carve_size(m * sizeof(cac_edge)) +
           carve_size(((size_t)j.t + 2) * 8) + carve_size(sizeof(EdgeSpan)) + carve_size(4);
  };
  size_t avail = 0;
  int rc = device_memory_available(ctx, &avail);
  if (rc) return rc;
  avail += ctx->lt_bytes;
  const size_t budget = (size_t)(0.35 * (double)avail);
  size_t b0 = 0;
  while (b0 < jobs.size()) {
    // the chunk: shared sort buffers for its largest graph, plus per-job scratch
    size_t nb = 0, per = 0, shared = 0;
    uint64_t mmax = 1, nmaxs = 1;
    size_t sort_tmp = 0, red_tmp = 0;
    while (b0 + nb < jobs.size()) {
      const SqrtJob& j = jobs[b0 + nb];
      const uint64_t m2 = std::max<uint64_t>(mmax, std::max<uint64_t>(j.m, 1));
      uint32_t bk = 1;
      while (bk < 32 && (1ull << bk) < j.t) ++bk;
      size_t st = 0, rt = 0;
      rocprim::double_buffer<uint64_t> kb(nullptr, nullptr), vb(nullptr, nullptr);
      CAC_HIP(ctx, rocprim::radix_sort_pairs(nullptr, st, kb, vb, m2, 0, 32 + bk, s));
      CAC_HIP(ctx, rocprim::reduce_by_key(nullptr, rt, (const uint64_t*)nullptr, (const uint64_t*)nullptr, m2,
                                          (uint64_t*)nullptr, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                          rocprim::plus<uint64_t>(), rocprim::equal_to<uint64_t>(), s));
      const size_t sh = 6 * carve_size(m2 * 8) + carve_size(std::max(std::max(st, rt), std::max(sort_tmp, red_tmp))) +
                        carve_size((scan_tiles(std::max<uint64_t>(nmaxs, j.n)) + 64) * 4) + 4096;
      const size_t add = job_bytes(j);
      if (nb > 0 && per + add + sh > budget) break;
      per += add;
      shared = sh;
      mmax = m2;
      nmaxs = std::max<uint64_t>(nmaxs, j.n);
      sort_tmp = std::max(sort_tmp, st);
      red_tmp = std::max(red_tmp, rt);
      ++nb;
    }
    rc = grow_buffer(ctx, &ctx->lt_buf, &ctx->lt_bytes, per + shared);
    if (rc) return rc;
    BufPin pin(ctx, kBufLt);  // the merged edge lists are the sparse roots
    Carver cv(ctx->lt_buf);
    uint64_t* k0 = cv.take<uint64_t>(mmax);
    uint64_t* k1 = cv.take<uint64_t>(mmax);
    uint64_t* v0 = cv.take<uint64_t>(mmax);
    uint64_t* v1 = cv.take<uint64_t>(mmax);
    uint64_t* ukeys = cv.take<uint64_t>(mmax);
    uint64_t* sums = cv.take<uint64_t>(mmax);
    void* tmp = cv.take<unsigned char>(std::max(sort_tmp, red_tmp));
    uint32_t* partial = cv.take<uint32_t>(scan_tiles(nmaxs) + 64);
    TrialScratch* d_scr = cv.take<TrialScratch>(nb);
    EdgeSpan* d_spans = cv.take<EdgeSpan>(nb);
    uint32_t* d_flags = cv.take<uint32_t>(nb);
    std::vector<TrialScratch> scr(nb);
    std::vector<uint32_t*> rank(nb);
    std::vector<cac_edge*> merged(nb);
    std::vector<unsigned long long*> deg(nb);
    for (size_t i = 0; i < nb; ++i) {
      const SqrtJob& j = jobs[b0 + i];
      TrialScratch& x = scr[i];
      x = TrialScratch{};
      x.edges = j.edges;
      x.m = j.m;
      x.n = j.n;
      x.t = j.t;
      x.trial = j.trial;
      x.slot = j.slot;
      x.lab = cv.take<uint32_t>(j.n);
      x.par = cv.take<uint32_t>(j.n);
      rank[i] = cv.take<uint32_t>(j.n);
      x.pre = cv.take<uint64_t>(std::max<uint64_t>(j.m, 1));
      x.draws = cv.take<uint2>(j.n);
      x.nv_out = cv.take<uint32_t>(16);
      merged[i] = cv.take<cac_edge>(std::max<uint64_t>(j.m, 1));
      deg[i] = cv.take<unsigned long long>((size_t)j.t + 2);  // [t] degrees, [t] count, [t + 1] min
    }
    rc = launch_mixed_flags(ctx, scr, d_spans, d_flags, s);
    if (rc) return rc;
    CAC_HIP(ctx, hipMemcpyAsync(d_scr, scr.data(), nb * sizeof(TrialScratch), hipMemcpyHostToDevice, s));
    prof_begin(ctx, "sample_trials", s);
    launch_sampling(ctx, nb, seed, d_scr, s);
    prof_end(ctx, s);
    CAC_LAUNCH_CHECK(ctx, "sample trials");
    std::vector<uint32_t> nv(nb);
    for (size_t i = 0; i < nb; ++i) CAC_HIP(ctx, hipMemcpyAsync(&nv[i], scr[i].nv_out, 4, hipMemcpyDeviceToHost, s));
    CAC_HIP(ctx, hipStreamSynchronize(s));
    std::vector<SparseRoot> roots;
    for (size_t i = 0; i < nb; ++i) {
      const SqrtJob& j = jobs[b0 + i];
      if (nv[i] > j.t || nv[i] < 2) {  // disconnected: a cut of weight 0 exists
        const int64_t zero = 0;
        CAC_HIP(ctx, hipMemcpyAsync(d_results + j.slot, &zero, 8, hipMemcpyHostToDevice, s));
        CAC_HIP(ctx, hipStreamSynchronize(s));
        continue;
      }
      const uint32_t k = nv[i];
      uint32_t bk = 1;
      while (bk < 32 && (1ull << bk) < j.t) ++bk;
      prof_begin(ctx, "contracted_edges", s);
      CAC_HIP(ctx, device_exclusive_scan<uint32_t>(LabLive{scr[i].lab}, LabRank{rank[i]}, j.n, partial,
                                                   (uint32_t*)nullptr, s));
      hipLaunchKernelGGL(k_lt_keys, dim3(grid_for(j.m, 256, 8192)), dim3(256), 0, s, j.edges, j.m,
                         (const uint32_t*)scr[i].lab, (const uint32_t*)rank[i], k0, v0);
      rocprim::double_buffer<uint64_t> kb(k0, k1), vb(v0, v1);
      size_t st = sort_tmp, rt = red_tmp;
      CAC_HIP(ctx, rocprim::radix_sort_pairs(tmp, st, kb, vb, j.m, 0, 32 + bk, s));
      uint64_t* d_count = (uint64_t*)(deg[i] + j.t);
      CAC_HIP(ctx, rocprim::reduce_by_key(tmp, rt, (const uint64_t*)kb.current(), (const uint64_t*)vb.current(), j.m,
                                          ukeys, sums, d_count, rocprim::plus<uint64_t>(),
                                          rocprim::equal_to<uint64_t>(), s));
      uint64_t count = 0;
      CAC_HIP(ctx, hipMemcpyAsync(&count, d_count, 8, hipMemcpyDeviceToHost, s));
      CAC_HIP(ctx, hipStreamSynchronize(s));
      uint64_t last = 0;
      if (count) {
        CAC_HIP(ctx, hipMemcpyAsync(&last, ukeys + count - 1, 8, hipMemcpyDeviceToHost, s));
        CAC_HIP(ctx, hipStreamSynchronize(s));
      }
      const uint64_t mm = count && last == ~0ull ? count - 1 : count;  // the loop key sorts last
      hipLaunchKernelGGL(k_lt_zero, dim3(grid_for((uint64_t)j.t + 2, 256, 1024)), dim3(256), 0, s, deg[i],
                         (uint64_t)j.t + 2);
      hipLaunchKernelGGL(k_lt_edges, dim3(grid_for(mm, 256, 8192)), dim3(256), 0, s, (const uint64_t*)ukeys,
                         (const uint64_t*)sums, mm, merged[i], deg[i]);
      unsigned long long* d_min = deg[i] + j.t + 1;
      const unsigned long long init = ~0ull;
      CAC_HIP(ctx, hipMemcpyAsync(d_min, &init, 8, hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(k_lt_min, dim3(grid_for(k, 256, 256)), dim3(256), 0, s, (const unsigned long long*)deg[i], k,
                         d_min);
      hipLaunchKernelGGL(k_lt_set, dim3(1), dim3(64), 0, s, d_results + j.slot, (const unsigned long long*)d_min);
      prof_end(ctx, s);
      CAC_LAUNCH_CHECK(ctx, "contracted edge list");
      unsigned long long bound = 0;
      CAC_HIP(ctx, hipMemcpyAsync(&bound, d_min, 8, hipMemcpyDeviceToHost, s));
      CAC_HIP(ctx, hipStreamSynchronize(s));
      if (mm > 0) roots.push_back({merged[i], mm, k, j.trial, j.slot, (int64_t)bound});
    }
    rc = ks_sparse_roots(ctx, roots, seed, d_results, s);
    if (rc) return rc;
    b0 += nb;
  }
  return CAC_OK;
}

int sqrt_cut_batch(cac_ctx* ctx, const std::vector<SqrtJob>& jobs, uint32_t seed, int64_t* d_results, hipStream_t s,
                   size_t budget) {
  if (jobs.empty()) return CAC_OK;
  bool any_large = false;
  for (const SqrtJob& j : jobs) any_large = any_large || j.t > kKsMaxS;
  if (any_large) {
    std::vector<SqrtJob> dense, large;
    for (const SqrtJob& j : jobs) (j.t <= kKsMaxS ? dense : large).push_back(j);
    const int rc = sqrt_large_batch(ctx, large, seed, d_results, s);
    if (rc) return rc;
    return sqrt_cut_batch(ctx, dense, seed, d_results, s, budget);
  }
  if (budget == 0) {
    size_t avail = 0;
    int rc0 = device_memory_available(ctx, &avail);
    if (rc0) return rc0;
    budget = std::max<size_t>((size_t)(0.45 * (double)avail), (size_t)1 << 28);
  }
  int rc = CAC_OK;
  size_t b0 = 0;
  while (b0 < jobs.size() && rc == CAC_OK) {
    size_t nb = 0, sample_bytes = 0, ks_bytes = 0;
    while (b0 + nb < jobs.size() && nb < kMaxJobsPerChunk) {
      const SqrtJob& j = jobs[b0 + nb];
      const size_t a = job_sample_bytes(j), k = job_ks_bytes(j);
      if (nb > 0 && sample_bytes + a + ks_bytes + k > budget) break;
      sample_bytes += a;
      ks_bytes += k;
      ++nb;
    }
    // whole waves of sampling workgroups: a last partial wave would leave most
    // of the GPU idle while it runs
    const size_t wave = sampling_wave(ctx);
    if (b0 + nb < jobs.size() && nb > wave) nb -= nb % wave;
    const size_t bytes = sample_bytes + carve_size(nb * sizeof(TrialScratch)) + 2 * carve_size(nb * 8) +
                         carve_size(nb * sizeof(EdgeSpan)) + carve_size(nb * 4) + 1024;
    rc = grow_buffer(ctx, &ctx->smp_buf, &ctx->smp_bytes, bytes);
    if (rc) return rc;
    BufPin pin(ctx, kBufSmp);  // the sampled matrices are the dense roots
    Carver cv(ctx->smp_buf);
    hipError_t e = hipSuccess;
    TrialScratch* d_scr = cv.take<TrialScratch>(nb);
    uint32_t* d_nv = cv.take<uint32_t>(nb);
    int64_t* d_bounds = cv.take<int64_t>(nb);
    EdgeSpan* d_spans = cv.take<EdgeSpan>(nb);
    uint32_t* d_flags = cv.take<uint32_t>(nb);
    std::vector<TrialScratch> scr(nb);
    for (size_t i = 0; i < nb; ++i) {
      const SqrtJob& j = jobs[b0 + i];
      TrialScratch& x = scr[i];
      x.edges = j.edges;
      x.m = j.m;
      x.n = j.n;
      x.t = j.t;
      x.trial = j.trial;
      x.slot = j.slot;
      x.lab = cv.take<uint32_t>(j.n);
      x.par = cv.take<uint32_t>(j.n);
      x.pre = cv.take<uint64_t>(std::max<uint64_t>(j.m, 1));
      x.draws = cv.take<uint2>(j.n);
      x.mat = cv.take<int64_t>((size_t)j.t * j.t);
      x.rsum = cv.take<int64_t>(j.t);
      x.nv_out = d_nv + i;
    }
    rc = launch_mixed_flags(ctx, scr, d_spans, d_flags, s);
    if (rc) return rc;
    e = hipMemcpyAsync(d_scr, scr.data(), nb * sizeof(TrialScratch), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) rc = hip_fail(ctx, e, "job descriptors");
    if (rc == CAC_OK) {
      prof_begin(ctx, "sample_trials", s);
      launch_sampling(ctx, nb, seed, d_scr, s);
      prof_end(ctx, s);
      prof_begin(ctx, "densify_trials", s);
      rc = densify_attr(ctx);
      if (rc == CAC_OK)
        hipLaunchKernelGGL(k_densify_trials, dim3((unsigned)nb), dim3(kSmpBlock), kDensifyLds, s,
                           (const TrialScratch*)d_scr, d_bounds, d_results);
      prof_end(ctx, s);
      e = hipGetLastError();
      if (e != hipSuccess) rc = hip_fail(ctx, e, "sampling kernels");
    }
    std::vector<uint32_t> nvs(nb);
    std::vector<int64_t> hb(nb);
    if (rc == CAC_OK) {
      (void)hipMemcpyAsync(nvs.data(), d_nv, nb * 4, hipMemcpyDeviceToHost, s);
      (void)hipMemcpyAsync(hb.data(), d_bounds, nb * 8, hipMemcpyDeviceToHost, s);
      e = hipStreamSynchronize(s);
      if (e != hipSuccess) rc = hip_fail(ctx, e, "sampling kernels");
    }
    if (rc == CAC_OK) {
      std::vector<KsRoot> roots;
      for (size_t i = 0; i < nb; ++i)
        if (nvs[i] >= 2 && nvs[i] <= scr[i].t)
          roots.push_back({scr[i].mat, scr[i].rsum, nvs[i], scr[i].trial, scr[i].slot, hb[i]});
      rc = ks_run(ctx, roots, seed, d_results, s);
    }
    (void)hipStreamSynchronize(s);
    b0 += nb;
  }
  return rc;
}

// Trials trial_ids of one graph; raw cut of trial_ids[i] to d_results[i].
int sqrt_cut_run(cac_ctx* ctx, const cac_edge* d_edges, uint64_t m, uint32_t n, uint32_t t, uint32_t seed,
                 const std::vector<uint32_t>& trial_ids, int64_t* d_results, hipStream_t s) {
  std::vector<SqrtJob> jobs(trial_ids.size());
  for (size_t i = 0; i < trial_ids.size(); ++i) jobs[i] = {d_edges, m, n, t, trial_ids[i], (uint32_t)i};
  return sqrt_cut_batch(ctx, jobs, seed, d_results, s);
}

}  // namespace cac

namespace cac {

// One trial's shrink phase for cac_sample_contract: labels (canonical) and the
// merged k x k matrix, k = super-vertices left (t unless disconnected).
int sample_contract_one(cac_ctx* ctx, const cac_edge* d_edges, uint64_t m, uint32_t n, uint32_t t, uint32_t seed,
                        uint32_t trial, SampleOut* out, hipStream_t s) {
  *out = SampleOut{};
  const size_t bytes = carve_size((size_t)n * 4) * 2 + carve_size(std::max<uint64_t>(m, 1) * 8) +
                       carve_size((size_t)n * 8) + 4 * carve_size(64) + carve_size(sizeof(TrialScratch)) +
                       carve_size(sizeof(EdgeSpan)) + carve_size(4);
  CAC_HIP(ctx, hipMalloc(&out->arena, bytes));
  Carver cv(out->arena);
  TrialScratch h{};
  h.edges = d_edges;
  h.m = m;
  h.n = n;
  h.t = t;
  h.trial = trial;
  h.slot = 0;
  TrialScratch* d_scr = cv.take<TrialScratch>(1);
  h.lab = cv.take<uint32_t>(n);
  h.par = cv.take<uint32_t>(n);
  h.pre = cv.take<uint64_t>(std::max<uint64_t>(m, 1));
  h.draws = cv.take<uint2>(n);
  h.nv_out = cv.take<uint32_t>(1);
  int64_t* d_misc = cv.take<int64_t>(2);
  EdgeSpan* d_span = cv.take<EdgeSpan>(1);
  uint32_t* d_flag = cv.take<uint32_t>(1);
  h.mat = nullptr;
  h.rsum = nullptr;
  std::vector<TrialScratch> one{h};
  int rc = launch_mixed_flags(ctx, one, d_span, d_flag, s);
  if (rc) return rc;
  h = one[0];
  CAC_HIP(ctx, hipMemcpyAsync(d_scr, &h, sizeof h, hipMemcpyHostToDevice, s));
  launch_sampling(ctx, 1, seed, d_scr, s);
  CAC_LAUNCH_CHECK(ctx, "sample trial");
  uint32_t k = 0;
  CAC_HIP(ctx, hipMemcpyAsync(&k, h.nv_out, 4, hipMemcpyDeviceToHost, s));
  CAC_HIP(ctx, hipStreamSynchronize(s));
  CAC_HIP(ctx, hipMalloc((void**)&h.mat, std::max<size_t>((size_t)k * k, 1) * 8));
  h.t = k;
  CAC_HIP(ctx, hipMemcpyAsync(d_scr, &h, sizeof h, hipMemcpyHostToDevice, s));
  rc = densify_attr(ctx);
  if (rc) return rc;
  hipLaunchKernelGGL(k_densify_trials, dim3(1), dim3(kSmpBlock), kDensifyLds, s, (const TrialScratch*)d_scr, d_misc,
                     d_misc + 1);
  CAC_LAUNCH_CHECK(ctx, "densify trial");
  CAC_HIP(ctx, hipStreamSynchronize(s));
  out->lab = h.lab;
  out->mat = h.mat;
  out->k = k;
  return CAC_OK;
}

void sample_out_free(SampleOut* o) {
  if (o->mat) (void)hipFree(o->mat);
  if (o->arena) (void)hipFree(o->arena);
  *o = SampleOut{};
}

}  // namespace cac
