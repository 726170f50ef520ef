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
//   1. prefix of the edge weights with loops masked to 0 (u64, block scan);
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
#include <vector>

#include "cac_common.h"
#include "cac_internal.h"
#include "cac_ks.h"
#include "cac_ks_kernels.h"

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
};

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

// Union-find over the first L draws from a clean forest; returns the hooks made.
__device__ uint32_t probe(const uint2* draws, uint32_t L, uint32_t* par, uint32_t n, uint64_t* lds) {
  for (uint32_t x = threadIdx.x; x < n; x += kSmpBlock) par[x] = x;
  block_sync_global();  // no wave may see the previous probe's forest
  uint64_t mine = 0;
  for (uint32_t d = threadIdx.x; d < L; d += kSmpBlock) mine += uf_union(par, draws[d].x, draws[d].y) ? 1 : 0;
  block_sync_global();  // the relabel reads roots after the last hook
  return (uint32_t)bsum64(mine, lds);
}

__global__ void __launch_bounds__(kSmpBlock) k_sample_trials(uint32_t seed, const TrialScratch* __restrict__ scratch) {
  __shared__ uint64_t lds[kSmpWaves + 2];
  const TrialScratch S = scratch[blockIdx.x];
  const cac_edge* __restrict__ edges = S.edges;
  const uint64_t m = S.m;
  const uint32_t n = S.n, t = S.t, trial = S.trial;
  for (uint32_t v = threadIdx.x; v < n; v += kSmpBlock) S.lab[v] = v;
  block_sync_global();
  uint32_t nv = n;
  for (uint32_t round = 0; nv > t; ++round) {
    // 1. masked inclusive prefix of the weights
    uint64_t carry = 0;
    constexpr int kItems = 4;
    for (uint64_t base = 0; base < m; base += (uint64_t)kSmpBlock * kItems) {
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
        if (e < m) S.pre[e] = run;
      }
      carry += tot;
      __syncthreads();
    }
    block_sync_global();  // the draws read the prefix written by every wave
    const uint64_t W = carry;
    if (W == 0) break;  // only loops left: more than t components
    // 2. draws
    const uint32_t D = nv;
    for (uint32_t d = threadIdx.x; d < D; d += kSmpBlock) {
      const u32x4 rnd = philox4x32(trial, round, d, 0u, seed, kDomSample);
      const uint64_t r = uniform1(lo64(rnd), W);
      uint64_t lo = 0, hi = m - 1;
      while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (S.pre[mid] >= r) hi = mid;
        else lo = mid + 1;
      }
      const cac_edge x = edges[lo];
      S.draws[d] = make_uint2(S.lab[x.u], S.lab[x.v]);
    }
    block_sync_global();
    // 3. shortest draw prefix that leaves t super-vertices
    uint32_t L = D;
    const uint32_t all = probe(S.draws, D, S.par, n, lds);
    if (nv - all <= t) {
      uint32_t lo = 1, hi = D;
      while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (nv - probe(S.draws, mid, S.par, n, lds) <= t) hi = mid;
        else lo = mid + 1;
      }
      L = lo;
    }
    const uint32_t hooks = probe(S.draws, L, S.par, n, lds);
    // 4. canonical relabel
    for (uint32_t v = threadIdx.x; v < n; v += kSmpBlock) S.lab[v] = uf_root(S.par, S.lab[v]);
    block_sync_global();
    nv -= hooks;
  }
  if (threadIdx.x == 0) *S.nv_out = nv;
}

// Rank the canonical labels, accumulate the surviving edges into the k x k
// matrix (k == t unless the sampling stopped early on a disconnected graph) and
// compute the trial's bound (minimum weighted degree).
// bounds is indexed by job (blockIdx.x), results by the job's slot.
__global__ void __launch_bounds__(kSmpBlock) k_densify_trials(const TrialScratch* __restrict__ scratch, int64_t* bounds,
                                                              int64_t* results) {
  __shared__ uint32_t wtot[kSmpWaves];
  __shared__ uint64_t lds[kSmpWaves + 2];
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
  block_sync_global();  // ranks and zeroed rows before the accumulation
  for (uint64_t e = threadIdx.x; e < m; e += kSmpBlock) {
    const cac_edge x = edges[e];
    const uint32_t a = rank[S.lab[x.u]], b = rank[S.lab[x.v]];
    if (a == b) continue;
    atomicAdd((unsigned long long*)&S.mat[(uint64_t)a * k + b], (unsigned long long)x.w);
    atomicAdd((unsigned long long*)&S.mat[(uint64_t)b * k + a], (unsigned long long)x.w);
  }
  block_sync_global();
  // row sums (the Karger-Stein capacities) and bound = min row sum: one wave per row
  uint64_t best = ~0ull;
  for (uint32_t r = threadIdx.x >> 6; r < k; r += kSmpWaves) {
    uint64_t acc = 0;
    for (uint32_t x = threadIdx.x & 63; x < k; x += 64) acc += (uint64_t)S.mat[(uint64_t)r * k + x];
    acc = wsum64(acc);
    if (S.rsum && (threadIdx.x & 63) == 0) S.rsum[r] = (int64_t)acc;
    best = acc < best ? acc : best;
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
int sqrt_cut_batch(cac_ctx* ctx, const std::vector<SqrtJob>& jobs, uint32_t seed, int64_t* d_results, hipStream_t s,
                   size_t budget) {
  if (jobs.empty()) return CAC_OK;
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
    const size_t bytes = sample_bytes + carve_size(nb * sizeof(TrialScratch)) + 2 * carve_size(nb * 8) + 1024;
    rc = grow_buffer(ctx, &ctx->smp_buf, &ctx->smp_bytes, bytes);
    if (rc) return rc;
    Carver cv(ctx->smp_buf);
    hipError_t e = hipSuccess;
    TrialScratch* d_scr = cv.take<TrialScratch>(nb);
    uint32_t* d_nv = cv.take<uint32_t>(nb);
    int64_t* d_bounds = cv.take<int64_t>(nb);
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
    e = hipMemcpyAsync(d_scr, scr.data(), nb * sizeof(TrialScratch), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) rc = hip_fail(ctx, e, "job descriptors");
    if (rc == CAC_OK) {
      prof_begin(ctx, "sample_trials", s);
      hipLaunchKernelGGL(k_sample_trials, dim3((unsigned)nb), dim3(kSmpBlock), 0, s, seed, (const TrialScratch*)d_scr);
      prof_end(ctx, s);
      prof_begin(ctx, "densify_trials", s);
      hipLaunchKernelGGL(k_densify_trials, dim3((unsigned)nb), dim3(kSmpBlock), 0, s, (const TrialScratch*)d_scr,
                         d_bounds, d_results);
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
                       carve_size((size_t)n * 8) + 4 * carve_size(64) + carve_size(sizeof(TrialScratch));
  CAC_HIP(ctx, hipMalloc(&out->arena, bytes));
  Carver cv(out->arena);
  TrialScratch h;
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
  h.mat = nullptr;
  h.rsum = nullptr;
  CAC_HIP(ctx, hipMemcpyAsync(d_scr, &h, sizeof h, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_sample_trials, dim3(1), dim3(kSmpBlock), 0, s, seed, (const TrialScratch*)d_scr);
  CAC_LAUNCH_CHECK(ctx, "sample trial");
  uint32_t k = 0;
  CAC_HIP(ctx, hipMemcpyAsync(&k, h.nv_out, 4, hipMemcpyDeviceToHost, s));
  CAC_HIP(ctx, hipStreamSynchronize(s));
  CAC_HIP(ctx, hipMalloc((void**)&h.mat, std::max<size_t>((size_t)k * k, 1) * 8));
  h.t = k;
  CAC_HIP(ctx, hipMemcpyAsync(d_scr, &h, sizeof h, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_densify_trials, dim3(1), dim3(kSmpBlock), 0, s, (const TrialScratch*)d_scr, d_misc,
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
