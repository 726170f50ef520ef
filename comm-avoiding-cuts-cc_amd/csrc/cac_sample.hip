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
#include <functional>
#include <mutex>
#include <unordered_map>
#include <vector>

#include <cstring>


#include "cac_common.h"
#include "cac_internal.h"
#include "cac_ks.h"
#include "cac_ks_kernels.h"
#include "cac_scan.h"
#include "cac_sort.h"
#include "cac_wave.h"
#include "cac_trial.h"

namespace cac {
namespace {

constexpr int kSmpBlock = 1024;
constexpr int kSmpWaves = kSmpBlock / 64;
constexpr int kSmpIndex = 4096;  // LDS segment index of the weight prefix (32 KiB)

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

__device__ inline uint64_t wsum64(uint64_t x) { return wave_sum<uint64_t>(x); }

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

// Root of x without path splitting: no entry but a hooked root's changes.
__device__ inline uint32_t uf_root_fixed(const uint32_t* par, uint32_t x) {
  uint32_t p = par[x];
  while (p != x) {
    x = p;
    p = par[x];
  }
  return x;
}

// Union without path splitting; on a hook, *hooked is the root that got a
// parent, so resetting it to itself undoes the hook exactly.
__device__ inline bool uf_union_fixed(uint32_t* par, uint32_t a, uint32_t b, uint32_t* hooked) {
  uint32_t ra = uf_root_fixed(par, a), rb = uf_root_fixed(par, b);
  while (ra != rb) {
    const uint32_t lo = min(ra, rb), hi = max(ra, rb);
    const uint32_t old = atomicCAS(par + hi, hi, lo);
    if (old == hi) {
      *hooked = hi;
      return true;
    }
    ra = uf_root_fixed(par, old);
    rb = uf_root_fixed(par, lo);
  }
  return false;
}

// Shortest prefix of the D draws whose unions make `need` hooks (all D make at
// least that many); leaves par holding exactly that prefix's forest and
// returns its hooks (== need).  From the round's starting forest the draws are
// united a chunk of kSmpBlock at a time (one per thread), with no path
// splitting, and a chunk that stays short of `need` is kept.  The chunk that
// reaches it is undone (every thread resets the root its hook changed) and
// the prefix length inside it is binary-searched the same way.  A prefix's
// hook count does not depend on the order of its unions, and each draw adds at
// most one hook, so this is the length the reference's draw-by-draw
// contraction stops at.
__device__ uint32_t prefix_search(const uint2* draws, uint32_t D, uint32_t need, uint32_t* par, uint64_t* lds) {
  for (uint32_t d = threadIdx.x; d < D; d += kSmpBlock) {  // back to the round's starting forest
    const uint2 x = draws[d];
    par[x.x] = x.x;
    par[x.y] = x.y;
  }
  block_sync_global();
  uint32_t total = 0;
  for (uint32_t c0 = 0; c0 < D; c0 += kSmpBlock) {
    const uint32_t len = min((uint32_t)kSmpBlock, D - c0);
    uint32_t mine = 0;
    bool hooked = false;
    if (threadIdx.x < len) {
      const uint2 x = draws[c0 + threadIdx.x];
      hooked = uf_union_fixed(par, x.x, x.y, &mine);
    }
    block_sync_global();
    const uint32_t ch = (uint32_t)bsum64(hooked ? 1 : 0, lds);
    if (total + ch < need) {  // the whole chunk stays
      total += ch;
      continue;
    }
    if (hooked) par[mine] = mine;  // undo the chunk
    block_sync_global();
    uint32_t lo = 1, hi = len;
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      uint32_t m2 = 0;
      bool h2 = false;
      if (threadIdx.x < mid) {
        const uint2 x = draws[c0 + threadIdx.x];
        h2 = uf_union_fixed(par, x.x, x.y, &m2);
      }
      block_sync_global();
      const uint32_t c2 = (uint32_t)bsum64(h2 ? 1 : 0, lds);
      if (h2) par[m2] = m2;
      block_sync_global();
      if (total + c2 >= need) hi = mid;
      else lo = mid + 1;
    }
    uint32_t m3 = 0;
    bool h3 = false;
    if (threadIdx.x < lo) {
      const uint2 x = draws[c0 + threadIdx.x];
      h3 = uf_union_fixed(par, x.x, x.y, &m3);
    }
    block_sync_global();
    return total + (uint32_t)bsum64(h3 ? 1 : 0, lds);
  }
  return total;
}

// kMinWaves = 8 holds the kernel to 64 VGPRs (8 waves per SIMD), so two 16-wave
// workgroups share a CU, at the price of a few spilled registers: for launches
// of more workgroups than CUs.  kMinWaves = 1 leaves the registers free, for
// small launches whose trials each have a CU to themselves.
//
// kLdsLab keeps the trial's labels in LDS as u16 (every vertex id below 2^16)
// and writes them to S.lab once at the end.  The prefix scan reads two labels
// per edge per round at random; with hundreds of trials in flight their label
// arrays do not fit in L2, so from global memory each such read fetches a whole
// line from HBM, many times the edge's own 16 bytes.
//
// Dynamic LDS: idx_cap u64 segment-index entries, then (kLdsLab) n u16 labels.
template <int kMinWaves, bool kLdsLab>
__global__ void __launch_bounds__(kSmpBlock, kMinWaves)
    k_sample_trials(const TrialScratch* __restrict__ scratch, uint32_t idx_cap) {
  __shared__ uint64_t lds[kSmpWaves + 2];
  extern __shared__ uint64_t smp_dyn[];
  // Every G-th prefix value (the last of each segment of G edges): a draw
  // finds its segment here, then searches only that segment in HBM.
  uint64_t* seg_last = smp_dyn;
  uint16_t* slab = reinterpret_cast<uint16_t*>(smp_dyn + idx_cap);
  const TrialScratch S = scratch[blockIdx.x];
  const cac_edge* __restrict__ edges = S.edges;
  const uint64_t m = S.m;
  uint32_t gshift = 0;  // segments of 2^gshift edges, at most idx_cap of them
  while (((m + (1ull << gshift) - 1) >> gshift) > (uint64_t)idx_cap) ++gshift;
  const uint32_t nseg = (uint32_t)((m + (1ull << gshift) - 1) >> gshift);
  const uint32_t n = S.n, t = S.t, trial = S.trial, seed = S.seed;
  auto lab_get = [&](uint32_t v) -> uint32_t {
    if constexpr (kLdsLab) return slab[v];
    else return S.lab[v];
  };
  auto lab_set = [&](uint32_t v, uint32_t x) {
    if constexpr (kLdsLab) slab[v] = (uint16_t)x;
    else S.lab[v] = x;
  };
  for (uint32_t v = threadIdx.x; v < n; v += kSmpBlock) {
    lab_set(v, v);
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
          if (lab_get(x.u) != lab_get(x.v)) w[k] = x.w;
        }
        s += w[k];
      }
      // block exclusive scan of s
      const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
      const uint64_t inc = wave_incl_scan<uint64_t>(s);
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
          if ((((e + 1) >> gshift) << gshift) == e + 1 || e + 1 == m) seg_last[e >> gshift] = run;
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
    const uint32_t D = S.dpr ? S.dpr : nv;
    for (uint32_t d = threadIdx.x; d < D; d += kSmpBlock) {
      const u32x4 rnd = philox4x32(trial, round, d, 0u, seed, kDomSample);
      const uint64_t r = uniform1(lo64(rnd), W);
      uint64_t lo = 0, hi = 0;
      if (flat) {
        lo = hi = (r - 1) / wu;
      } else {  // the first segment whose last prefix reaches r, then inside it
        uint32_t a = 0, b = nseg - 1;
        while (a < b) {
          const uint32_t mid = (a + b) >> 1;
          if (seg_last[mid] >= r) b = mid;
          else a = mid + 1;
        }
        lo = (uint64_t)a << gshift;
        hi = min(lo + (1ull << gshift), m) - 1;
      }
      while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if ((narrow ? (uint64_t)pre32[mid] : S.pre[mid]) >= r) hi = mid;
        else lo = mid + 1;
      }
      const cac_edge x = edges[lo];
      S.draws[d] = make_uint2(lab_get(x.u), lab_get(x.v));
    }
    block_sync_global();
    // 3. shortest draw prefix that leaves t super-vertices
    uint32_t hooks = probe(S.draws, D, 0, S.par, lds);
    if (nv - hooks <= t) hooks = prefix_search(S.draws, D, nv - t, S.par, lds);
    // else all D draws are kept, and the forest of the first probe stays
    // 4. canonical relabel
    for (uint32_t v = threadIdx.x; v < n; v += kSmpBlock) lab_set(v, uf_root(S.par, lab_get(v)));
    block_sync_global();
    nv -= hooks;
  }
  if constexpr (kLdsLab) {  // each thread copies the labels it relabelled last
    for (uint32_t v = threadIdx.x; v < n; v += kSmpBlock) S.lab[v] = slab[v];
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
// kRankLabels: the labels are replaced by their ranks before the accumulation,
// so an edge endpoint costs one dependent load instead of two (label, then
// rank); for callers that do not read the labels afterwards.
template <bool kRankLabels>
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
    const uint32_t inc = wave_incl_scan<uint32_t>(live);
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
  if constexpr (kRankLabels) {
    block_sync_global();  // every rank written
    for (uint32_t x = threadIdx.x; x < n; x += kSmpBlock) S.lab[x] = rank[S.lab[x]];
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
    if constexpr (!kRankLabels) {
#pragma unroll
      for (int j = 0; j < kDensifyEdges; ++j) {
        a[j] = rank[a[j]];
        b[j] = rank[b[j]];
      }
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
  best = wave_min<uint64_t>(best);
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

// Sort keys of one job's edges: (job << 2 bk) | (a << bk) | b of the ranked
// endpoints a < b, the low 2 bk bits all ones for a loop (dropped after the
// sort); value w.
__global__ void __launch_bounds__(256) k_lt_keys(const cac_edge* __restrict__ edges, uint64_t m,
                                                 const uint32_t* __restrict__ lab, const uint32_t* __restrict__ rank,
                                                 uint64_t job_hi, uint32_t bk, uint64_t* __restrict__ keys,
                                                 uint64_t* __restrict__ vals) {
  const uint64_t loop = (1ull << (2 * bk)) - 1;
  for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < m; e += (uint64_t)gridDim.x * 256) {
    const cac_edge x = edges[e];
    const uint32_t a = rank[lab[x.u]], b = rank[lab[x.v]];
    keys[e] = job_hi | (a == b ? loop : ((uint64_t)min(a, b) << bk) | max(a, b));
    vals[e] = x.w;
  }
}

// Merged (key, weight) pairs to cac_edge records in place (the 16-byte layouts
// coincide: the key's 8 bytes become u, v), and every job's weighted degrees.
__global__ void __launch_bounds__(256) k_lt_edges(uint64_t* __restrict__ pairs, uint64_t count, uint32_t bk,
                                                  unsigned long long* const* __restrict__ deg) {
  const uint64_t low = (1ull << bk) - 1;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (uint64_t)gridDim.x * 256) {
    const uint64_t k = pairs[2 * i], w = pairs[2 * i + 1];
    const uint32_t a = (uint32_t)((k >> bk) & low), b = (uint32_t)(k & low);
    reinterpret_cast<cac_edge*>(pairs)[i] = cac_edge{a, b, w};
    unsigned long long* d = deg[k >> (2 * bk)];
    atomicAdd(&d[a], (unsigned long long)w);
    atomicAdd(&d[b], (unsigned long long)w);
  }
}

// firsts[j] = head_idx[offsets[j]]: where each job's merged list starts.
__global__ void k_lt_firsts(const uint32_t* __restrict__ head_idx, const uint64_t* __restrict__ offsets, uint32_t count,
                            uint32_t* __restrict__ firsts) {
  for (uint32_t j = blockIdx.x * 256 + threadIdx.x; j < count; j += gridDim.x * 256) firsts[j] = head_idx[offsets[j]];
}

__global__ void __launch_bounds__(256) k_fill_u64(uint64_t* p, uint64_t count, uint64_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (uint64_t)gridDim.x * 256) p[i] = v;
}

__global__ void __launch_bounds__(256) k_lt_zero(unsigned long long* p, uint64_t count) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (uint64_t)gridDim.x * 256) p[i] = 0;
}

// *out = min(deg[0..k)) (out preset to ~0), and results[slot] = that bound.
__global__ void __launch_bounds__(256) k_lt_min(const unsigned long long* deg, uint32_t k, unsigned long long* out) {
  __shared__ unsigned long long w[4];
  unsigned long long b = ~0ull;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < k; i += (uint64_t)gridDim.x * 256) b = min(b, deg[i]);
  b = (unsigned long long)wave_min<uint64_t>((uint64_t)b);
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
    err = hipFuncSetAttribute((const void*)k_densify_trials<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kDensifyLds);
    if (err == hipSuccess)
      err = hipFuncSetAttribute((const void*)k_densify_trials<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)kDensifyLds);
  });
  CAC_HIP(ctx, err);
  return CAC_OK;
}


// Dynamic LDS a sampling workgroup may take: the CU's 160 KiB less the
// kernel's static arrays, with room to spare.
constexpr size_t kSmpLdsMax = 160 * 1024 - 1024;

// Launch shape of k_sample_trials for trials of at most nmax vertices.
struct SmpShape {
  bool lds_lab;      // labels in LDS as u16 (nmax <= 2^16)
  uint32_t idx_cap;  // segment-index entries
  size_t dyn;        // dynamic LDS bytes
};

SmpShape sampling_shape(uint64_t nmax) {
  SmpShape sh{};
  sh.lds_lab = nmax <= 65536u;
  const size_t lab_bytes = sh.lds_lab ? (((size_t)nmax * 2 + 15) & ~size_t(15)) : 0;
  sh.idx_cap = (uint32_t)std::min<size_t>(kSmpIndex, ((kSmpLdsMax - lab_bytes) / 8) & ~size_t(63));
  sh.dyn = (size_t)sh.idx_cap * 8 + lab_bytes;
  return sh;
}

// The sampling kernels' dynamic LDS exceeds the default limit: raise it once.
int sampling_attr(cac_ctx* ctx) {
  static std::once_flag once;
  static hipError_t err = hipSuccess;
  std::call_once(once, [] {
    const void* fns[] = {(const void*)k_sample_trials<8, false>, (const void*)k_sample_trials<1, false>,
                         (const void*)k_sample_trials<8, true>, (const void*)k_sample_trials<1, true>};
    for (const void* f : fns) {
      const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSmpLdsMax);
      if (e != hipSuccess && err == hipSuccess) err = e;
    }
  });
  CAC_HIP(ctx, err);
  return CAC_OK;
}

// Sampling workgroups the device runs at once (workgroups per CU times CUs).
size_t sampling_wave(cac_ctx* ctx) {
  if (ctx->smp_wave == 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_sample_trials<8, false>, kSmpBlock,
                                                     (size_t)kSmpIndex * 8) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess) {
      (void)hipGetLastError();
      per_cu = 1;
      cus = 1;
    }
    ctx->smp_wave = (size_t)std::max(per_cu, 1) * (size_t)std::max(cus, 1);
  }
  return ctx->smp_wave;
}


// One sampling launch of nb trials of at most nmax vertices and mmax edges, in
// the form suited to its size: a launch of more trials than CUs runs two
// 1024-thread workgroups per CU (64 VGPRs, some spilled) when their LDS fits
// twice in a CU and the context's tuning table does not say one for mmax; a
// smaller launch gives each trial a CU of its own with the registers free.
// The draws do not depend on the form.
int launch_sampling(cac_ctx* ctx, size_t nb, uint64_t mmax, uint64_t nmax, const TrialScratch* d_scr,
                    hipStream_t s) {
  int rc = sampling_attr(ctx);
  if (rc) return rc;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess) {
    (void)hipGetLastError();
    cus = 1;
  }
  const SmpShape sh = sampling_shape(nmax);
  const int cls = mmax < (1ull << 16) ? 0 : mmax < (1ull << 20) ? 1 : mmax < (1ull << 24) ? 2 : 3;
  const int per_cu = ctx->sample_occupancy[cls] ? ctx->sample_occupancy[cls] : 2;
  const bool two = nb > (size_t)cus && per_cu == 2 && 2 * (sh.dyn + 1024) <= 160 * 1024;
  const dim3 grid((unsigned)nb), block(kSmpBlock);
  if (sh.lds_lab) {
    if (two) hipLaunchKernelGGL((k_sample_trials<8, true>), grid, block, sh.dyn, s, d_scr, sh.idx_cap);
    else hipLaunchKernelGGL((k_sample_trials<1, true>), grid, block, sh.dyn, s, d_scr, sh.idx_cap);
  } else {
    if (two) hipLaunchKernelGGL((k_sample_trials<8, false>), grid, block, sh.dyn, s, d_scr, sh.idx_cap);
    else hipLaunchKernelGGL((k_sample_trials<1, false>), grid, block, sh.dyn, s, d_scr, sh.idx_cap);
  }
  return CAC_OK;
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

// Draws per round of the high-concurrency rule: (unsigned) pow((float) n0,
// 1 + epsilon / 2) with epsilon = 0.1f, in float like
// WeightedIteratedSparseSampling::edgesToSamplePerProcessor
// (src/WeightedIteratedSparseSampling.hpp:66-68, src/IteratedSparseSampling.hpp:38).
uint32_t hc_draws_host(uint32_t n0) {
  const float epsilon = 0.1f;
  return (uint32_t)std::pow((float)n0, 1 + epsilon / 2);
}

// Draws per round of a job under the context's sampling rule (0: |V|).
static uint32_t job_dpr(const cac_ctx* ctx, uint32_t n) {
  return ctx->sample_rule == CAC_SAMPLE_HC ? hc_draws_host(n) : 0u;
}

// Device bytes one sampling job needs beside the Karger-Stein levels (dpr:
// draws per round, 0 = |V|).
size_t job_sample_bytes(const SqrtJob& j, uint32_t dpr) {
  return carve_size((size_t)j.n * 4) * 2 + carve_size(std::max<uint64_t>(j.m, 1) * 8) +
         carve_size((size_t)std::max(j.n, dpr) * 8) + carve_size((size_t)j.t * j.t * 8) + carve_size((size_t)j.t * 8) +
         256;
}
// Upper estimate of the Karger-Stein level arenas of one root of size t: each
// level holds about t^2 entries in copy-on-write work rows plus the contracted
// children, and two levels are alive at once.
size_t job_ks_bytes(const SqrtJob& j) { return ks_live_bytes_per_root(j.t); }

// Edge-count bound from which a job of the batch runs in the wide form.  One
// workgroup scans a trial's edges in time proportional to m on one CU, while
// the workgroup-form jobs keep every CU busy for about (their sum of m) / CUs;
// a trial above that is the launch's straggler.  From the largest m down, a job
// goes wide (if m reaches the context's threshold) while it would still be a
// straggler beside the jobs left to the workgroup form, which spend fewer
// CU-seconds per edge and so take the rest.
static uint64_t wide_cut(cac_ctx* ctx, const std::vector<SqrtJob>& jobs) {
  const uint64_t wmin = wide_min_edges(ctx);
  if (wmin == ~0ull) return wmin;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cus < 1) {
    (void)hipGetLastError();
    cus = 1;
  }
  std::vector<uint64_t> ms;
  uint64_t narrow_total = 0;
  for (const SqrtJob& j : jobs) {
    ms.push_back(j.m);
    narrow_total += j.m;
  }
  std::sort(ms.begin(), ms.end(), std::greater<uint64_t>());
  uint64_t cut = ~0ull;
  for (const uint64_t m : ms) {
    if (m < wmin || (long double)m * cus <= (long double)narrow_total) break;
    cut = m;
    narrow_total -= m;
  }
  return cut;
}

// Trials with t > 8192, in chunks sized to the free memory: one sampling
// launch for the chunk; then the contracted edge lists of all its trials at
// once -- endpoints ranked per trial, keys (trial, a, b), one radix sort of
// every trial's keys together, one segmented sum of parallel weights, the
// merged lists and their weighted degrees -- with two host synchronisations
// per chunk (the super-vertex counts, then the list ranges and bounds); then
// one sparse Karger-Stein pass over all of the chunk's trials (node 1 each).
// Raw cut of each job to d_results[slot].
static int sqrt_large_batch(cac_ctx* ctx, const std::vector<SqrtJob>& jobs_in, int64_t* d_results,
                            hipStream_t s) {
  // trials of large edge lists sample in the wide form: they go last in every
  // chunk, after the workgroup-per-trial ones
  const uint64_t wmin = wide_cut(ctx, jobs_in);
  std::vector<SqrtJob> jobs;
  for (const SqrtJob& j : jobs_in)
    if (j.m < wmin) jobs.push_back(j);
  for (const SqrtJob& j : jobs_in)
    if (j.m >= wmin) jobs.push_back(j);
  auto job_bytes = [ctx, wmin](const SqrtJob& j) {
    const uint64_t m = std::max<uint64_t>(j.m, 1);
    return carve_size(sizeof(TrialScratch)) + 3 * carve_size((size_t)j.n * 4) + carve_size(m * 8) +
           carve_size((size_t)std::max(j.n, job_dpr(ctx, j.n)) * 8) + carve_size(64) + carve_size(((size_t)j.t + 2) * 8) +
           carve_size(sizeof(EdgeSpan)) + carve_size(4) + 64 * m +  // 64 m: the shared merge buffers
           (j.m >= wmin ? wide_trial_bytes(j.m, j.n) : 0);
  };
  size_t avail = 0;
  int rc = device_memory_available(ctx, &avail);
  if (rc) return rc;
  avail += ctx->lt_bytes;
  const size_t budget = (size_t)(0.35 * (double)avail);
  size_t b0 = 0;
  while (b0 < jobs.size()) {
    size_t nb = 0, per = 0;
    uint64_t N = 0, nmax = 1;
    uint32_t tmax = 2;
    while (b0 + nb < jobs.size()) {
      const SqrtJob& j = jobs[b0 + nb];
      const size_t add = job_bytes(j);
      if (nb > 0 && (per + add > budget || N + j.m >= 0xFFFFFFF0ull || nb + 1 > 4096)) break;
      per += add;
      N += j.m;
      nmax = std::max<uint64_t>(nmax, j.n);
      tmax = std::max(tmax, j.t);
      ++nb;
    }
    uint32_t bk = 1;
    while (bk < 32 && (1ull << bk) < tmax) ++bk;
    uint32_t jb = 0;
    while ((1ull << jb) < nb) ++jb;
    const int key_bits = (int)(2 * bk + jb);
    if (key_bits > 64) return fail(ctx, CAC_EINVAL, "contracted graph too large for 64-bit merge keys");
    const uint64_t Np = (std::max<uint64_t>(N, 1) + 31) & ~uint64_t(31);
    size_t w0 = 0, wide_bytes = 0;  // jobs [b0, b0 + w0) run one workgroup each, the rest wide
    for (size_t i = 0; i < nb; ++i) {
      const SqrtJob& j = jobs[b0 + i];
      if (j.m < wmin) ++w0;
      else wide_bytes += wide_trial_bytes(j.m, j.n);
    }
    const size_t shared = carve_size(Np * 16) * 2 + carve_size((Np + 1) * 4) + carve_size(radix_sort_temp_bytes(Np)) +
                          carve_size(segmented_sum_temp_bytes(Np)) + carve_size((scan_tiles(nmax) + 64) * 4) +
                          carve_size((nb + 1) * 8) + carve_size((nb + 1) * 4) + carve_size(nb * 8) +
                          carve_size(wide_bytes) + 4096;
    rc = grow_buffer(ctx, &ctx->lt_buf, &ctx->lt_bytes, per + shared);
    if (rc) return rc;
    BufPin pin(ctx, kBufLt);  // the merged edge lists are the sparse roots
    Carver cv(ctx->lt_buf);
    uint64_t* buf_a = cv.take<uint64_t>(2 * Np);  // keys | values, then free for the merged pairs
    uint64_t* buf_b = cv.take<uint64_t>(2 * Np);
    uint32_t* head_idx = cv.take<uint32_t>(Np + 1);
    void* sort_tmp = cv.take<unsigned char>(radix_sort_temp_bytes(Np));
    void* seg_tmp = cv.take<unsigned char>(segmented_sum_temp_bytes(Np));
    uint32_t* partial = cv.take<uint32_t>(scan_tiles(nmax) + 64);
    uint64_t* d_offsets = cv.take<uint64_t>(nb + 1);
    uint32_t* d_firsts = cv.take<uint32_t>(nb + 1);
    unsigned long long** d_deg = cv.take<unsigned long long*>(nb);
    TrialScratch* d_scr = cv.take<TrialScratch>(nb);
    EdgeSpan* d_spans = cv.take<EdgeSpan>(nb);
    uint32_t* d_flags = cv.take<uint32_t>(nb);
    void* wide_work = cv.take<unsigned char>(std::max<size_t>(wide_bytes, 1));
    std::vector<TrialScratch> scr(nb);
    std::vector<uint32_t*> rank(nb);
    std::vector<unsigned long long*> deg(nb);
    std::vector<uint64_t> offsets(nb + 1, 0);
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
      x.seed = j.seed;
      x.dpr = job_dpr(ctx, j.n);
      x.lab = cv.take<uint32_t>(j.n);
      x.par = cv.take<uint32_t>(j.n);
      rank[i] = cv.take<uint32_t>(j.n);
      x.pre = cv.take<uint64_t>(std::max<uint64_t>(j.m, 1));
      x.draws = cv.take<uint2>(std::max(j.n, x.dpr));
      x.nv_out = cv.take<uint32_t>(16);
      deg[i] = cv.take<unsigned long long>((size_t)j.t + 2);  // [t] degrees, [t + 1] their minimum
      offsets[i + 1] = offsets[i] + j.m;
    }
    std::vector<TrialScratch> nar(scr.begin(), scr.begin() + w0), wid(scr.begin() + w0, scr.end());
    rc = launch_mixed_flags(ctx, nar, d_spans, d_flags, s);
    if (rc) return rc;
    for (size_t i = 0; i < w0; ++i) scr[i] = nar[i];
    CAC_HIP(ctx, hipMemcpyAsync(d_scr, scr.data(), nb * sizeof(TrialScratch), hipMemcpyHostToDevice, s));
    CAC_HIP(ctx, hipMemcpyAsync(d_offsets, offsets.data(), (nb + 1) * 8, hipMemcpyHostToDevice, s));
    CAC_HIP(ctx, hipMemcpyAsync(d_deg, deg.data(), nb * sizeof(void*), hipMemcpyHostToDevice, s));
    prof_begin(ctx, "sample_trials", s);
    if (w0 > 0) {
      uint64_t mmax_jobs = 0, nmax_jobs = 0;
      for (size_t i = 0; i < w0; ++i) {
        mmax_jobs = std::max<uint64_t>(mmax_jobs, jobs[b0 + i].m);
        nmax_jobs = std::max<uint64_t>(nmax_jobs, jobs[b0 + i].n);
      }
      rc = launch_sampling(ctx, w0, mmax_jobs, nmax_jobs, d_scr, s);
      if (rc) return rc;
    }
    if (w0 < nb) {
      rc = wide_sample(ctx, wid, d_scr + w0, wide_work, s);
      if (rc) return rc;
    }
    prof_end(ctx, s);
    CAC_LAUNCH_CHECK(ctx, "sample trials");
    std::vector<uint32_t> nv(nb);
    for (size_t i = 0; i < nb; ++i) CAC_HIP(ctx, hipMemcpyAsync(&nv[i], scr[i].nv_out, 4, hipMemcpyDeviceToHost, s));
    CAC_HIP(ctx, hipStreamSynchronize(s));
    // keys of every job (a disconnected trial's edges all become loops: its
    // cut of 0 needs no list)
    prof_begin(ctx, "contracted_edges", s);
    uint64_t* keys = buf_a;
    uint64_t* vals = buf_a + Np;
    for (size_t i = 0; i < nb; ++i) {
      const SqrtJob& j = jobs[b0 + i];
      const bool disconnected = nv[i] > j.t || nv[i] < 2;
      if (!disconnected)
        CAC_HIP(ctx, device_exclusive_scan<uint32_t>(LabLive{scr[i].lab}, LabRank{rank[i]}, j.n, partial,
                                                     (uint32_t*)nullptr, s));
      if (j.m == 0) continue;
      if (disconnected) {  // all loops
        hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(j.m, 256, 8192)), dim3(256), 0, s, keys + offsets[i], j.m,
                           ((uint64_t)i << (2 * bk)) | ((1ull << (2 * bk)) - 1));
        continue;
      }
      hipLaunchKernelGGL(k_lt_keys, dim3(grid_for(j.m, 256, 8192)), dim3(256), 0, s, j.edges, j.m,
                         (const uint32_t*)scr[i].lab, (const uint32_t*)rank[i], (uint64_t)i << (2 * bk), bk,
                         keys + offsets[i], vals + offsets[i]);
    }
    CAC_LAUNCH_CHECK(ctx, "merge keys");
    bool in_alt = false;
    rc = radix_sort_pairs(ctx, keys, vals, buf_b, buf_b + Np, N, key_bits, sort_tmp, &in_alt, s);
    if (rc) return rc;
    uint64_t* sk = in_alt ? buf_b : buf_a;
    uint64_t* pairs = in_alt ? buf_a : buf_b;  // the free buffer takes the merged pairs
    rc = segmented_sum(ctx, sk, sk + Np, N, (1ull << (2 * bk)) - 1, pairs, pairs + 1, 2, head_idx, seg_tmp, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_lt_firsts, dim3(grid_for(nb + 1, 256, 64)), dim3(256), 0, s, (const uint32_t*)head_idx,
                       (const uint64_t*)d_offsets, (uint32_t)(nb + 1), d_firsts);
    for (size_t i = 0; i < nb; ++i)
      hipLaunchKernelGGL(k_lt_zero, dim3(grid_for((uint64_t)jobs[b0 + i].t + 2, 256, 1024)), dim3(256), 0, s, deg[i],
                         (uint64_t)jobs[b0 + i].t + 2);
    std::vector<uint32_t> firsts(nb + 1);
    CAC_HIP(ctx, hipMemcpyAsync(firsts.data(), d_firsts, (nb + 1) * 4, hipMemcpyDeviceToHost, s));
    CAC_HIP(ctx, hipStreamSynchronize(s));
    const uint64_t merged = firsts[nb];
    if (merged)
      hipLaunchKernelGGL(k_lt_edges, dim3(grid_for(merged, 256, 8192)), dim3(256), 0, s, pairs, merged, bk,
                         (unsigned long long* const*)d_deg);
    std::vector<unsigned long long> bounds(nb, 0);
    for (size_t i = 0; i < nb; ++i) {
      const SqrtJob& j = jobs[b0 + i];
      if (nv[i] > j.t || nv[i] < 2) {  // disconnected: a cut of weight 0 exists
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, (uint64_t*)(d_results + j.slot), (uint64_t)1, 0ull);
        continue;
      }
      unsigned long long* d_min = deg[i] + j.t + 1;
      hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, (uint64_t*)d_min, (uint64_t)1, ~0ull);
      hipLaunchKernelGGL(k_lt_min, dim3(grid_for(nv[i], 256, 256)), dim3(256), 0, s,
                         (const unsigned long long*)deg[i], nv[i], d_min);
      hipLaunchKernelGGL(k_lt_set, dim3(1), dim3(64), 0, s, d_results + j.slot, (const unsigned long long*)d_min);
      CAC_HIP(ctx, hipMemcpyAsync(&bounds[i], d_min, 8, hipMemcpyDeviceToHost, s));
    }
    prof_end(ctx, s);
    CAC_LAUNCH_CHECK(ctx, "contracted edge lists");
    CAC_HIP(ctx, hipStreamSynchronize(s));
    std::vector<SparseRoot> roots;
    for (size_t i = 0; i < nb; ++i) {
      const SqrtJob& j = jobs[b0 + i];
      const uint64_t mm = firsts[i + 1] - firsts[i];
      if (nv[i] > j.t || nv[i] < 2 || mm == 0) continue;
      roots.push_back({reinterpret_cast<cac_edge*>(pairs) + firsts[i], mm, nv[i], j.trial, j.slot,
                       (int64_t)bounds[i], j.seed});
    }
    rc = ks_sparse_roots(ctx, roots, d_results, s);
    if (rc) return rc;
    b0 += nb;
  }
  return CAC_OK;
}

// Dense-engine jobs (t <= 8192) in chunks sized to the budget: sampling,
// densification, then one breadth-first Karger-Stein pass per chunk, whatever
// graphs the chunk's jobs belong to.  wide: every job runs in the wide form
// (its rounds as grid-wide kernels), else one workgroup per job.
static int sqrt_dense_chunks(cac_ctx* ctx, const std::vector<SqrtJob>& jobs, int64_t* d_results, hipStream_t s,
                             size_t budget, bool wide) {
  int rc = CAC_OK;
  size_t b0 = 0;
  while (b0 < jobs.size() && rc == CAC_OK) {
    size_t nb = 0, sample_bytes = 0, ks_bytes = 0, wide_bytes = 0;
    const size_t cap = wide ? 4096 : kMaxJobsPerChunk;
    while (b0 + nb < jobs.size() && nb < cap) {
      const SqrtJob& j = jobs[b0 + nb];
      const size_t a = job_sample_bytes(j, job_dpr(ctx, j.n)), k = job_ks_bytes(j);
      const size_t wb = wide ? wide_trial_bytes(j.m, j.n) : 0;
      if (nb > 0 && sample_bytes + a + ks_bytes + k + wide_bytes + wb > budget) break;
      sample_bytes += a;
      ks_bytes += k;
      wide_bytes += wb;
      ++nb;
    }
    // whole waves of sampling workgroups: a last partial wave would leave most
    // of the GPU idle while it runs
    const size_t wave = sampling_wave(ctx);
    if (!wide && b0 + nb < jobs.size() && nb > wave) nb -= nb % wave;
    if (wide) {
      wide_bytes = 0;
      for (size_t i = 0; i < nb; ++i)
        wide_bytes += wide_trial_bytes(jobs[b0 + i].m, jobs[b0 + i].n);
    }
    const size_t bytes = sample_bytes + carve_size(nb * sizeof(TrialScratch)) + 2 * carve_size(nb * 8) +
                         carve_size(nb * sizeof(EdgeSpan)) + carve_size(nb * 4) + carve_size(wide_bytes) + 1024;
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
    void* wide_work = wide ? (void*)cv.take<unsigned char>(wide_bytes) : nullptr;
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
      x.seed = j.seed;
      x.dpr = job_dpr(ctx, j.n);
      x.lab = cv.take<uint32_t>(j.n);
      x.par = cv.take<uint32_t>(j.n);
      x.pre = cv.take<uint64_t>(std::max<uint64_t>(j.m, 1));
      x.draws = cv.take<uint2>(std::max(j.n, x.dpr));
      x.mat = cv.take<int64_t>((size_t)j.t * j.t);
      x.rsum = cv.take<int64_t>(j.t);
      x.nv_out = d_nv + i;
    }
    if (!wide) {
      rc = launch_mixed_flags(ctx, scr, d_spans, d_flags, s);
      if (rc) return rc;
    }
    e = hipMemcpyAsync(d_scr, scr.data(), nb * sizeof(TrialScratch), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) rc = hip_fail(ctx, e, "job descriptors");
    if (rc == CAC_OK && wide) {
      prof_begin(ctx, "wide_sample", s);
      rc = wide_sample(ctx, scr, d_scr, wide_work, s);
      prof_end(ctx, s);
      prof_begin(ctx, "wide_densify", s);
      if (rc == CAC_OK) rc = wide_densify(ctx, scr, d_scr, wide_work, d_bounds, d_results, s);
      prof_end(ctx, s);
    } else if (rc == CAC_OK) {
      prof_begin(ctx, "sample_trials", s);
      uint64_t mmax_jobs = 0, nmax_jobs = 0;
      for (size_t i = 0; i < nb; ++i) {
        mmax_jobs = std::max<uint64_t>(mmax_jobs, jobs[b0 + i].m);
        nmax_jobs = std::max<uint64_t>(nmax_jobs, jobs[b0 + i].n);
      }
      rc = launch_sampling(ctx, nb, mmax_jobs, nmax_jobs, d_scr, s);
      prof_end(ctx, s);
      prof_begin(ctx, "densify_trials", s);
      if (rc == CAC_OK) rc = densify_attr(ctx);
      if (rc == CAC_OK)
        hipLaunchKernelGGL(k_densify_trials<true>, dim3((unsigned)nb), dim3(kSmpBlock), kDensifyLds, s,
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
          roots.push_back({scr[i].mat, scr[i].rsum, nvs[i], scr[i].trial, scr[i].slot, hb[i], scr[i].seed});
      rc = ks_run(ctx, roots, d_results, s);
    }
    (void)hipStreamSynchronize(s);
    b0 += nb;
  }
  return rc;
}

// Runs sqrt-cut jobs (graph, trial) in chunks sized to the free device memory.
// Raw cut of job i goes to d_results[jobs[i].slot].
int sqrt_cut_batch(cac_ctx* ctx, const std::vector<SqrtJob>& jobs, int64_t* d_results, hipStream_t s,
                   size_t budget, bool by_size) {
  if (jobs.empty()) return CAC_OK;
  bool any_large = false;
  for (const SqrtJob& j : jobs) any_large = any_large || j.t > kKsMaxS;
  if (any_large) {
    std::vector<SqrtJob> dense, large;
    for (const SqrtJob& j : jobs) (j.t <= kKsMaxS ? dense : large).push_back(j);
    const int rc = sqrt_large_batch(ctx, large, d_results, s);
    if (rc) return rc;
    return sqrt_cut_batch(ctx, dense, d_results, s, budget, by_size);
  }
  if (budget == 0) {
    size_t avail = 0;
    int rc0 = device_memory_available(ctx, &avail);
    if (rc0) return rc0;
    budget = std::max<size_t>((size_t)(0.45 * (double)avail), (size_t)1 << 28);
  }
  // trials of large edge lists run in the wide form, the rest one workgroup each
  const uint64_t wmin = wide_cut(ctx, jobs);
  std::vector<SqrtJob> narrow, wide;
  for (const SqrtJob& j : jobs) (j.m >= wmin ? wide : narrow).push_back(j);
  // by_size: chunks of trials of like size when no workgroup-form trial reaches
  // the wide form's floor.  A chunk of small graphs then has a small label
  // array, its densify launch holds matrices of like size and its Karger-Stein
  // levels children of like size; results go by slot.  Above the floor the
  // wide cut has counted on the large trials sharing chunks with small ones.
  if (by_size) {
    const uint64_t floor_m = wide_min_edges(ctx);
    bool below_floor = true;
    for (const SqrtJob& j : narrow) below_floor = below_floor && j.m < floor_m;
    if (below_floor)
      std::stable_sort(narrow.begin(), narrow.end(), [](const SqrtJob& a, const SqrtJob& b) { return a.n > b.n; });
  }
  int rc = sqrt_dense_chunks(ctx, wide, d_results, s, budget, true);
  if (rc) return rc;
  return sqrt_dense_chunks(ctx, narrow, d_results, s, budget, false);
}

// Trials trial_ids of one graph; raw cut of trial_ids[i] to d_results[i].
int sqrt_cut_run(cac_ctx* ctx, const cac_edge* d_edges, uint64_t m, uint32_t n, uint32_t t, uint32_t seed,
                 const std::vector<uint32_t>& trial_ids, int64_t* d_results, hipStream_t s) {
  std::vector<SqrtJob> jobs(trial_ids.size());
  for (size_t i = 0; i < trial_ids.size(); ++i) jobs[i] = {d_edges, m, n, t, trial_ids[i], (uint32_t)i, seed};
  return sqrt_cut_batch(ctx, jobs, d_results, s);
}

}  // namespace cac

namespace cac {

// One trial's shrink phase for cac_sample_contract: labels (canonical) and the
// merged k x k matrix, k = super-vertices left (t unless disconnected).
int sample_contract_one(cac_ctx* ctx, const cac_edge* d_edges, uint64_t m, uint32_t n, uint32_t t, uint32_t seed,
                        uint32_t trial, SampleOut* out, hipStream_t s) {
  *out = SampleOut{};
  const uint32_t dpr = job_dpr(ctx, n);
  const size_t bytes = carve_size((size_t)n * 4) * 2 + carve_size(std::max<uint64_t>(m, 1) * 8) +
                       carve_size((size_t)std::max(n, dpr) * 8) + 4 * carve_size(64) + carve_size(sizeof(TrialScratch)) +
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
  h.seed = seed;
  h.dpr = dpr;
  TrialScratch* d_scr = cv.take<TrialScratch>(1);
  h.lab = cv.take<uint32_t>(n);
  h.par = cv.take<uint32_t>(n);
  h.pre = cv.take<uint64_t>(std::max<uint64_t>(m, 1));
  h.draws = cv.take<uint2>(std::max(n, dpr));
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
  rc = launch_sampling(ctx, 1, m, n, d_scr, s);
  if (rc) return rc;
  CAC_LAUNCH_CHECK(ctx, "sample trial");
  uint32_t k = 0;
  CAC_HIP(ctx, hipMemcpyAsync(&k, h.nv_out, 4, hipMemcpyDeviceToHost, s));
  CAC_HIP(ctx, hipStreamSynchronize(s));
  CAC_HIP(ctx, hipMalloc((void**)&h.mat, std::max<size_t>((size_t)k * k, 1) * 8));
  h.t = k;
  CAC_HIP(ctx, hipMemcpyAsync(d_scr, &h, sizeof h, hipMemcpyHostToDevice, s));
  rc = densify_attr(ctx);
  if (rc) return rc;
  hipLaunchKernelGGL(k_densify_trials<false>, dim3(1), dim3(kSmpBlock), kDensifyLds, s, (const TrialScratch*)d_scr, d_misc,
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
