// cac_wide.hip -- wide sqrt-cut trials on gfx950: one trial's sampling rounds
// and its densification as grid-wide kernels over its edge list.
//
// The workgroup-per-trial kernels (cac_sample.hip) scan a trial's whole edge
// list once per round with one CU.  A query of 10^7 edges has a dozen trials,
// so such a launch holds a dozen workgroups on 256 CUs and each runs for
// hundreds of milliseconds.  Here every round is a short sequence of kernels
// whose grids cover (edge tiles, trials), so one trial uses the whole GPU.
//
// Same algorithm as SequentialSquareRootCut::iteratedSampling
// (src/SequentialSquareRootCut.cpp:4-67) and the device kernel k_sample_trials,
// step for step, so draws, draw-prefix lengths and labels are bit-identical:
//   1. masked inclusive u64 prefix of the weights: per-tile totals, one scan of
//      the totals per trial, then the prefix written tile by tile;
//   2. D draws (|V|, or the high-concurrency count): Philox(ctr = {trial, round,
//      draw, 0}, key = {seed, 'SAMP'}), r = 1 + floor(x W / 2^64), edge = first
//      prefix >= r (search over tile totals, then inside the tile);
//   3. union-find over the draws.  When all D draws leave more than t
//      super-vertices they are all kept; otherwise the shortest draw prefix
//      that reaches t is binary-searched.  The search state lives on the
//      device per trial, so a probe step is four launches with no host wait;
//   4. canonical relabel (hooks attach the larger root under the smaller, so
//      every root is its set's minimum id).
// The host waits twice per round: after the first probe (does any trial
// search?) and after the relabel (which trials go on).
//
// Union-find here is touched by workgroups on different XCDs, whose L2s are
// not coherent with each other for plain loads within a kernel.  Parent reads
// and path-splitting writes are relaxed agent-scope atomics, so a find always
// sees the hooks other XCDs made and the CAS retry loop makes progress.
#include <algorithm>
#include <vector>

#include "cac_common.h"
#include "cac_internal.h"
#include "cac_ks.h"
#include "cac_trial.h"

namespace cac {
namespace {

constexpr int kWB = 256;                    // threads per block
constexpr int kWWaves = kWB / 64;
constexpr int kWItems = 16;                 // consecutive edges per thread in a tile
constexpr uint32_t kWTile = kWB * kWItems;  // edges (or vertices) per tile
constexpr int kWCtl = 64;                   // threads per block of the per-trial control kernels

enum : uint32_t { kWIdle = 0, kWKeep = 1, kWSearch = 2, kWFinal = 3, kWDone = 4 };

// A wide trial's control words (device).
struct WideState {
  uint64_t* tsum;    // [tiles] inclusive prefix of the tile totals
  uint64_t W;        // the round's masked total
  uint64_t bmin;     // densify: minimum row sum
  uint32_t nv;       // super-vertices
  uint32_t active;   // 1 while the trial is sampling this round
  uint32_t D;        // draws of the round
  uint32_t mode;     // kW* probe state of the round
  uint32_t lo, hi;   // draw-prefix search interval
  uint32_t last;     // length of the last probe (its forest is in par)
  uint32_t cur;      // length of this step's probe (0: none)
  uint32_t undo;     // draws whose endpoints this step resets first
  uint32_t hooks;    // hooks of the last probe
  uint32_t round;
  uint32_t pad;
};

__device__ inline uint32_t par_load(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void par_store(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline uint32_t wuf_root(uint32_t* par, uint32_t x) {
  uint32_t p = par_load(par + x);
  while (p != x) {
    const uint32_t g = par_load(par + p);
    if (g != p) par_store(par + x, g);  // path splitting: any ancestor is a valid parent
    x = p;
    p = g;
  }
  return x;
}

// Hooks the sets of a and b; true when this call joined two roots.
__device__ inline bool wuf_union(uint32_t* par, uint32_t a, uint32_t b) {
  uint32_t ra = wuf_root(par, a), rb = wuf_root(par, b);
  while (ra != rb) {
    const uint32_t lo = min(ra, rb), hi = max(ra, rb);
    const uint32_t old = atomicCAS(par + hi, hi, lo);
    if (old == hi) return true;
    ra = wuf_root(par, old);
    rb = wuf_root(par, lo);
  }
  return false;
}

__device__ inline uint64_t wave_sum64(uint64_t x) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

// Block-wide sum (every thread gets it).
__device__ inline uint64_t block_sum64(uint64_t x) {
  __shared__ uint64_t part[kWWaves];
  x = wave_sum64(x);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = x;
  __syncthreads();
  uint64_t t = 0;
#pragma unroll
  for (int w = 0; w < kWWaves; ++w) t += part[w];
  __syncthreads();
  return t;
}

// Exclusive block scan of x (every thread gets its offset and the total).
__device__ inline uint64_t block_excl64(uint64_t x, uint64_t* total) {
  __shared__ uint64_t part[kWWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t inc = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t y = __shfl_up(inc, d, 64);
    if (lane >= d) inc += y;
  }
  if (lane == 63) part[wave] = inc;
  __syncthreads();
  uint64_t off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kWWaves; ++w) {
    if (w < wave) off += part[w];
    tot += part[w];
  }
  __syncthreads();
  *total = tot;
  return off + inc - x;
}

// ---- sampling ---------------------------------------------------------------

// Grid (vertex blocks, trials).
__global__ void __launch_bounds__(kWB) k_w_init(const TrialScratch* __restrict__ scr, WideState* st) {
  const TrialScratch& S = scr[blockIdx.y];
  const uint32_t v = blockIdx.x * kWB + threadIdx.x;
  if (v < S.n) {
    S.lab[v] = v;
    S.par[v] = v;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    WideState& w = st[blockIdx.y];
    w.nv = S.n;
    w.active = S.n > S.t ? 1u : 0u;
    w.round = 0;
    w.mode = kWIdle;
    w.W = 0;
  }
}

// Masked weight of an edge under the trial's current labels.
__device__ inline uint64_t masked_w(const TrialScratch& S, uint64_t e) {
  const cac_edge x = S.edges[e];
  return S.lab[x.u] != S.lab[x.v] ? x.w : 0ull;
}

// Grid (edge tiles, trials): per-tile masked totals (vertices = false) or
// per-tile counts of canonical labels (vertices = true, for the ranks).
template <bool kVertices>
__global__ void __launch_bounds__(kWB) k_w_tile_sums(const TrialScratch* __restrict__ scr, WideState* st) {
  const TrialScratch& S = scr[blockIdx.y];
  WideState& w = st[blockIdx.y];
  if (!kVertices && !w.active) return;
  const uint64_t count = kVertices ? (uint64_t)S.n : S.m;
  const uint64_t base = (uint64_t)blockIdx.x * kWTile;
  if (base >= count) return;
  uint64_t s = 0;
#pragma unroll 4
  for (int j = 0; j < kWItems; ++j) {
    const uint64_t e = base + (uint64_t)j * kWB + threadIdx.x;
    if (e < count) s += kVertices ? (S.lab[e] == (uint32_t)e ? 1ull : 0ull) : masked_w(S, e);
  }
  s = block_sum64(s);
  if (threadIdx.x == 0) w.tsum[blockIdx.x] = s;
}

// Grid (trials), 1024 threads: inclusive scan of a trial's tile totals in
// place.  For the sampling scan it also sets the round's total, draw count and
// first probe (all D draws); a trial with nothing left to draw stops.
template <bool kVertices>
__global__ void __launch_bounds__(1024) k_w_tile_scan(const TrialScratch* __restrict__ scr, WideState* st) {
  const TrialScratch& S = scr[blockIdx.x];
  WideState& w = st[blockIdx.x];
  if (!kVertices && !w.active) return;
  const uint64_t count = kVertices ? (uint64_t)S.n : S.m;
  const uint32_t tiles = (uint32_t)((count + kWTile - 1) / kWTile);
  __shared__ uint64_t part[16];
  __shared__ uint64_t carry_s;
  if (threadIdx.x == 0) carry_s = 0;
  __syncthreads();
  for (uint32_t base = 0; base < tiles; base += 1024) {
    const uint32_t i = base + threadIdx.x;
    const uint64_t x = i < tiles ? w.tsum[i] : 0ull;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t inc = x;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t y = __shfl_up(inc, d, 64);
      if (lane >= d) inc += y;
    }
    if (lane == 63) part[wave] = inc;
    __syncthreads();
    uint64_t off = carry_s, tot = 0;
    for (int v = 0; v < 16; ++v) {
      if (v < wave) off += part[v];
      tot += part[v];
    }
    if (i < tiles) w.tsum[i] = off + inc;
    __syncthreads();
    if (threadIdx.x == 0) carry_s += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (kVertices) {
      w.bmin = ~0ull;
    } else {
      w.W = carry_s;
      if (carry_s == 0) {  // only loops left: more than t components
        w.active = 0;
      } else {
        w.D = S.dpr ? S.dpr : w.nv;
        w.mode = kWIdle;
        w.cur = w.D;
        w.undo = 0;
        w.hooks = 0;
      }
    }
  }
}

// Grid (edge tiles, trials): the masked inclusive prefix of the tile, offset by
// the totals of the tiles before it.  Each thread takes kWItems consecutive
// edges, so a wave reads and writes one contiguous run of the tile.
__global__ void __launch_bounds__(kWB) k_w_prefix(const TrialScratch* __restrict__ scr, WideState* st) {
  const TrialScratch& S = scr[blockIdx.y];
  const WideState& w = st[blockIdx.y];
  if (!w.active) return;
  const uint64_t base = (uint64_t)blockIdx.x * kWTile;
  if (base >= S.m) return;
  const uint64_t e0 = base + (uint64_t)threadIdx.x * kWItems;
  uint64_t v[kWItems], s = 0;
#pragma unroll
  for (int j = 0; j < kWItems; ++j) {
    v[j] = e0 + j < S.m ? masked_w(S, e0 + j) : 0ull;
    s += v[j];
  }
  uint64_t tot = 0;
  uint64_t run = block_excl64(s, &tot) + (blockIdx.x ? w.tsum[blockIdx.x - 1] : 0ull);
#pragma unroll
  for (int j = 0; j < kWItems; ++j) {
    run += v[j];
    if (e0 + j < S.m) S.pre[e0 + j] = run;
  }
}

// Grid (draw blocks, trials): the round's draws.
__global__ void __launch_bounds__(kWB) k_w_draws(const TrialScratch* __restrict__ scr, WideState* st) {
  const TrialScratch& S = scr[blockIdx.y];
  const WideState& w = st[blockIdx.y];
  if (!w.active) return;
  const uint32_t d = blockIdx.x * kWB + threadIdx.x;
  if (d >= w.D) return;
  const u32x4 rnd = philox4x32(S.trial, w.round, d, 0u, S.seed, kDomSample);
  const uint64_t r = uniform1(lo64(rnd), w.W);
  const uint32_t tiles = (uint32_t)((S.m + kWTile - 1) / kWTile);
  uint32_t a = 0, b = tiles - 1;  // first tile whose inclusive total reaches r
  while (a < b) {
    const uint32_t mid = (a + b) >> 1;
    if (w.tsum[mid] >= r) b = mid;
    else a = mid + 1;
  }
  uint64_t lo = (uint64_t)a * kWTile, hi = min(lo + kWTile, S.m) - 1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (S.pre[mid] >= r) hi = mid;
    else lo = mid + 1;
  }
  const cac_edge x = S.edges[lo];
  S.draws[d] = make_uint2(S.lab[x.u], S.lab[x.v]);
}

// Grid (draw blocks, trials): reset the endpoints of the first `undo` draws of
// the previous probe, so every label they touched is its own root again.
__global__ void __launch_bounds__(kWB) k_w_undo(const TrialScratch* __restrict__ scr, const WideState* st) {
  const TrialScratch& S = scr[blockIdx.y];
  const uint32_t d = blockIdx.x * kWB + threadIdx.x;
  if (d >= st[blockIdx.y].undo) return;
  const uint2 x = S.draws[d];
  par_store(S.par + x.x, x.x);
  par_store(S.par + x.y, x.y);
}

// Grid (draw blocks, trials): union over the first `cur` draws, counting hooks.
__global__ void __launch_bounds__(kWB) k_w_union(const TrialScratch* __restrict__ scr, WideState* st) {
  const TrialScratch& S = scr[blockIdx.y];
  WideState& w = st[blockIdx.y];
  const uint32_t L = w.cur;
  if ((uint64_t)blockIdx.x * kWB >= L) return;
  const uint32_t d = blockIdx.x * kWB + threadIdx.x;
  uint64_t h = 0;
  if (d < L) {
    const uint2 x = S.draws[d];
    h = wuf_union(S.par, x.x, x.y) ? 1 : 0;
  }
  h = block_sum64(h);
  if (threadIdx.x == 0 && h) atomicAdd(&w.hooks, (uint32_t)h);
}

// Per trial, after the probe of all D draws: keep them all, or search.
__global__ void k_w_after_first(const TrialScratch* __restrict__ scr, WideState* st, uint32_t k) {
  const uint32_t i = blockIdx.x * kWCtl + threadIdx.x;
  if (i >= k) return;
  WideState& w = st[i];
  if (!w.active) return;
  w.cur = 0;
  w.undo = 0;
  if (w.nv - w.hooks > scr[i].t) {
    w.mode = kWKeep;
  } else {
    w.mode = kWSearch;
    w.lo = 1;
    w.hi = w.D;
    w.last = w.D;
  }
}

// Per trial, before a search step: the step's probe length and reset length.
__global__ void k_w_step_begin(WideState* st, uint32_t k) {
  const uint32_t i = blockIdx.x * kWCtl + threadIdx.x;
  if (i >= k) return;
  WideState& w = st[i];
  w.cur = 0;
  w.undo = 0;
  if (!w.active || w.mode != kWSearch) return;
  if (w.lo < w.hi) {
    w.cur = w.lo + (w.hi - w.lo) / 2;
  } else if (w.last != w.lo) {
    w.cur = w.lo;
    w.mode = kWFinal;
  } else {
    w.mode = kWDone;  // the last probe was at the final length: its forest stays
    return;
  }
  w.undo = w.last;
  w.hooks = 0;
}

// Per trial, after a search step's probe.
__global__ void k_w_step_end(const TrialScratch* __restrict__ scr, WideState* st, uint32_t k) {
  const uint32_t i = blockIdx.x * kWCtl + threadIdx.x;
  if (i >= k) return;
  WideState& w = st[i];
  if (!w.active || w.cur == 0) return;
  w.last = w.cur;
  if (w.mode == kWFinal) {
    w.mode = kWDone;
  } else if (w.nv - w.hooks <= scr[i].t) {
    w.hi = w.cur;
  } else {
    w.lo = w.cur + 1;
  }
  w.cur = 0;
  w.undo = 0;
}

// Grid (vertex blocks, trials): canonical relabel.
__global__ void __launch_bounds__(kWB) k_w_relabel(const TrialScratch* __restrict__ scr, const WideState* st) {
  const TrialScratch& S = scr[blockIdx.y];
  if (!st[blockIdx.y].active) return;
  const uint32_t v = blockIdx.x * kWB + threadIdx.x;
  if (v < S.n) S.lab[v] = wuf_root(S.par, S.lab[v]);
}

__global__ void k_w_round_end(const TrialScratch* __restrict__ scr, WideState* st, uint32_t k) {
  const uint32_t i = blockIdx.x * kWCtl + threadIdx.x;
  if (i >= k) return;
  WideState& w = st[i];
  if (!w.active || (w.mode != kWKeep && w.mode != kWDone)) return;  // the host sees an unfinished search
  w.nv -= w.hooks;
  w.round += 1;
  w.active = w.nv > scr[i].t ? 1u : 0u;
  w.mode = kWIdle;
}

__global__ void k_w_nv_out(const TrialScratch* __restrict__ scr, const WideState* st, uint32_t k) {
  const uint32_t i = blockIdx.x * kWCtl + threadIdx.x;
  if (i < k) *scr[i].nv_out = st[i].nv;
}

// ---- densification ----------------------------------------------------------

// Grid (vertex tiles, trials): rank of each canonical label (into par).
__global__ void __launch_bounds__(kWB) k_w_rank(const TrialScratch* __restrict__ scr, const WideState* st) {
  const TrialScratch& S = scr[blockIdx.y];
  const WideState& w = st[blockIdx.y];
  if (w.nv > S.t) return;
  const uint64_t base = (uint64_t)blockIdx.x * kWTile;
  if (base >= S.n) return;
  const uint64_t v0 = base + (uint64_t)threadIdx.x * kWItems;
  uint32_t live[kWItems];
  uint64_t s = 0;
#pragma unroll
  for (int j = 0; j < kWItems; ++j) {
    live[j] = v0 + j < S.n && S.lab[v0 + j] == (uint32_t)(v0 + j) ? 1u : 0u;
    s += live[j];
  }
  uint64_t tot = 0;
  uint64_t run = block_excl64(s, &tot) + (blockIdx.x ? w.tsum[blockIdx.x - 1] : 0ull);
#pragma unroll
  for (int j = 0; j < kWItems; ++j) {
    if (live[j]) S.par[v0 + j] = (uint32_t)run;
    run += live[j];
  }
}

// Grid (matrix blocks, trials): zero the nv x nv matrix.
__global__ void __launch_bounds__(kWB) k_w_zero(const TrialScratch* __restrict__ scr, const WideState* st) {
  const TrialScratch& S = scr[blockIdx.y];
  const uint32_t k = st[blockIdx.y].nv;
  if (k > S.t) return;
  const uint64_t cells = (uint64_t)k * k;
  for (uint64_t i = (uint64_t)blockIdx.x * kWB + threadIdx.x; i < cells; i += (uint64_t)gridDim.x * kWB) S.mat[i] = 0;
}

// Grid (edge tiles, trials): accumulate the surviving edges into the matrix.
__global__ void __launch_bounds__(kWB) k_w_accum(const TrialScratch* __restrict__ scr, const WideState* st) {
  const TrialScratch& S = scr[blockIdx.y];
  const uint32_t k = st[blockIdx.y].nv;
  if (k > S.t) return;
  const uint64_t base = (uint64_t)blockIdx.x * kWTile;
  if (base >= S.m) return;
#pragma unroll 4
  for (int j = 0; j < kWItems; ++j) {
    const uint64_t e = base + (uint64_t)j * kWB + threadIdx.x;
    if (e >= S.m) break;
    const cac_edge x = S.edges[e];
    const uint32_t la = S.lab[x.u], lb = S.lab[x.v];
    if (la == lb) continue;
    const uint32_t a = S.par[la], b = S.par[lb];
    atomicAdd((unsigned long long*)&S.mat[(uint64_t)a * k + b], (unsigned long long)x.w);
    atomicAdd((unsigned long long*)&S.mat[(uint64_t)b * k + a], (unsigned long long)x.w);
  }
}

// Grid (row groups, trials), one wave per row: row sums and their minimum.
__global__ void __launch_bounds__(kWB) k_w_rowsum(const TrialScratch* __restrict__ scr, WideState* st) {
  const TrialScratch& S = scr[blockIdx.y];
  WideState& w = st[blockIdx.y];
  const uint32_t k = w.nv;
  if (k > S.t) return;
  const uint32_t r = blockIdx.x * kWWaves + (threadIdx.x >> 6);
  if (r >= k) return;
  uint64_t acc = 0;
  for (uint32_t x = threadIdx.x & 63; x < k; x += 64) acc += (uint64_t)S.mat[(uint64_t)r * k + x];
  acc = wave_sum64(acc);
  if ((threadIdx.x & 63) == 0) {
    if (S.rsum) S.rsum[r] = (int64_t)acc;
    atomicMin((unsigned long long*)&w.bmin, (unsigned long long)acc);
  }
}

__global__ void k_w_bounds(const TrialScratch* __restrict__ scr, const WideState* st, uint32_t k, int64_t* bounds,
                           int64_t* results) {
  const uint32_t i = blockIdx.x * kWCtl + threadIdx.x;
  if (i >= k) return;
  const int64_t b = st[i].nv > scr[i].t ? 0 : (int64_t)st[i].bmin;  // disconnected: a cut of weight 0
  bounds[i] = b;
  results[scr[i].slot] = b;
}

unsigned grid_y(size_t k) { return (unsigned)k; }

}  // namespace

uint64_t wide_min_edges(const cac_ctx* ctx) {
  const int32_t v = ctx->wide_log2;
  if (v == 0) return 1ull << 20;
  if (v >= 63) return ~0ull;
  return 1ull << v;
}

size_t wide_trial_bytes(uint64_t m, uint32_t n) {
  const uint64_t tiles = (std::max<uint64_t>(m, n) + kWTile - 1) / kWTile + 1;
  return carve_size(tiles * 8) + carve_size(sizeof(WideState));
}

namespace {

// The per-trial state array and tile buffers, carved from work in order.
WideState* wide_state(const std::vector<TrialScratch>& h, void* work, std::vector<WideState>* hs) {
  Carver cv(work);
  WideState* d_st = cv.take<WideState>(h.size());
  hs->assign(h.size(), WideState{});
  for (size_t i = 0; i < h.size(); ++i) {
    const uint64_t tiles = (std::max<uint64_t>(h[i].m, h[i].n) + kWTile - 1) / kWTile + 1;
    (*hs)[i].tsum = cv.take<uint64_t>(tiles);
  }
  return d_st;
}

}  // namespace

int wide_sample(cac_ctx* ctx, const std::vector<TrialScratch>& h, const TrialScratch* d_scr, void* work,
                hipStream_t s) {
  const size_t k = h.size();
  if (k == 0) return CAC_OK;
  if (k > 65535) return fail(ctx, CAC_EINVAL, "too many wide trials in one launch");
  std::vector<WideState> hs;
  WideState* d_st = wide_state(h, work, &hs);
  uint64_t mmax = 1, nmax = 1;
  for (const TrialScratch& x : h) {
    mmax = std::max<uint64_t>(mmax, x.m);
    nmax = std::max<uint64_t>(nmax, x.n);
  }
  CAC_HIP(ctx, hipMemcpyAsync(d_st, hs.data(), k * sizeof(WideState), hipMemcpyHostToDevice, s));
  const unsigned ctl = (unsigned)((k + kWCtl - 1) / kWCtl);
  const dim3 vgrid((unsigned)((nmax + kWB - 1) / kWB), grid_y(k));
  const dim3 egrid((unsigned)((mmax + kWTile - 1) / kWTile), grid_y(k));
  hipLaunchKernelGGL(k_w_init, vgrid, dim3(kWB), 0, s, d_scr, d_st);
  CAC_LAUNCH_CHECK(ctx, "wide sampling init");
  std::vector<WideState> now(k);
  CAC_HIP(ctx, hipMemcpyAsync(now.data(), d_st, k * sizeof(WideState), hipMemcpyDeviceToHost, s));
  CAC_HIP(ctx, hipStreamSynchronize(s));
  for (;;) {
    uint64_t dmax = 0;
    for (size_t i = 0; i < k; ++i)
      if (now[i].active) dmax = std::max<uint64_t>(dmax, h[i].dpr ? h[i].dpr : now[i].nv);
    if (dmax == 0) break;
    const dim3 dgrid((unsigned)((dmax + kWB - 1) / kWB), grid_y(k));
    hipLaunchKernelGGL(k_w_tile_sums<false>, egrid, dim3(kWB), 0, s, d_scr, d_st);
    hipLaunchKernelGGL(k_w_tile_scan<false>, dim3(grid_y(k)), dim3(1024), 0, s, d_scr, d_st);
    hipLaunchKernelGGL(k_w_prefix, egrid, dim3(kWB), 0, s, d_scr, d_st);
    hipLaunchKernelGGL(k_w_draws, dgrid, dim3(kWB), 0, s, d_scr, d_st);
    hipLaunchKernelGGL(k_w_union, dgrid, dim3(kWB), 0, s, d_scr, d_st);
    hipLaunchKernelGGL(k_w_after_first, dim3(ctl), dim3(kWCtl), 0, s, d_scr, d_st, (uint32_t)k);
    CAC_LAUNCH_CHECK(ctx, "wide sampling round");
    CAC_HIP(ctx, hipMemcpyAsync(now.data(), d_st, k * sizeof(WideState), hipMemcpyDeviceToHost, s));
    CAC_HIP(ctx, hipStreamSynchronize(s));
    bool search = false;
    for (size_t i = 0; i < k; ++i) search = search || (now[i].active && now[i].mode == kWSearch);
    if (search) {
      // a search over [1, D] ends within ceil(log2 D) probes, plus the final one
      int steps = 2;
      while ((1ull << (steps - 2)) < dmax) ++steps;
      for (int it = 0; it < steps; ++it) {
        hipLaunchKernelGGL(k_w_step_begin, dim3(ctl), dim3(kWCtl), 0, s, d_st, (uint32_t)k);
        hipLaunchKernelGGL(k_w_undo, dgrid, dim3(kWB), 0, s, d_scr, (const WideState*)d_st);
        hipLaunchKernelGGL(k_w_union, dgrid, dim3(kWB), 0, s, d_scr, d_st);
        hipLaunchKernelGGL(k_w_step_end, dim3(ctl), dim3(kWCtl), 0, s, d_scr, d_st, (uint32_t)k);
      }
      CAC_LAUNCH_CHECK(ctx, "wide sampling search");
    }
    hipLaunchKernelGGL(k_w_relabel, vgrid, dim3(kWB), 0, s, d_scr, (const WideState*)d_st);
    hipLaunchKernelGGL(k_w_round_end, dim3(ctl), dim3(kWCtl), 0, s, d_scr, d_st, (uint32_t)k);
    CAC_LAUNCH_CHECK(ctx, "wide sampling relabel");
    CAC_HIP(ctx, hipMemcpyAsync(now.data(), d_st, k * sizeof(WideState), hipMemcpyDeviceToHost, s));
    CAC_HIP(ctx, hipStreamSynchronize(s));
    for (size_t i = 0; i < k; ++i)
      if (now[i].mode != kWIdle) return fail(ctx, CAC_EDEVICE, "wide sampling: draw-prefix search did not end");
  }
  hipLaunchKernelGGL(k_w_nv_out, dim3(ctl), dim3(kWCtl), 0, s, d_scr, (const WideState*)d_st, (uint32_t)k);
  CAC_LAUNCH_CHECK(ctx, "wide sampling result");
  return CAC_OK;
}

int wide_densify(cac_ctx* ctx, const std::vector<TrialScratch>& h, const TrialScratch* d_scr, void* work,
                 int64_t* bounds, int64_t* results, hipStream_t s) {
  const size_t k = h.size();
  if (k == 0) return CAC_OK;
  std::vector<WideState> hs;
  WideState* d_st = wide_state(h, work, &hs);
  std::vector<WideState> now(k);
  CAC_HIP(ctx, hipMemcpyAsync(now.data(), d_st, k * sizeof(WideState), hipMemcpyDeviceToHost, s));
  CAC_HIP(ctx, hipStreamSynchronize(s));
  uint64_t mmax = 1, nmax = 1, kmax = 1;
  for (size_t i = 0; i < k; ++i) {
    mmax = std::max<uint64_t>(mmax, h[i].m);
    nmax = std::max<uint64_t>(nmax, h[i].n);
    if (now[i].nv <= h[i].t) kmax = std::max<uint64_t>(kmax, now[i].nv);
  }
  const unsigned ctl = (unsigned)((k + kWCtl - 1) / kWCtl);
  const dim3 tgrid((unsigned)((nmax + kWTile - 1) / kWTile), grid_y(k));
  const dim3 egrid((unsigned)((mmax + kWTile - 1) / kWTile), grid_y(k));
  const uint64_t cells = kmax * kmax;
  const dim3 zgrid((unsigned)std::min<uint64_t>((cells + kWB * 8 - 1) / (kWB * 8), 65536), grid_y(k));
  const dim3 rgrid((unsigned)((kmax + kWWaves - 1) / kWWaves), grid_y(k));
  hipLaunchKernelGGL(k_w_tile_sums<true>, tgrid, dim3(kWB), 0, s, d_scr, d_st);
  hipLaunchKernelGGL(k_w_tile_scan<true>, dim3(grid_y(k)), dim3(1024), 0, s, d_scr, d_st);
  hipLaunchKernelGGL(k_w_rank, tgrid, dim3(kWB), 0, s, d_scr, (const WideState*)d_st);
  hipLaunchKernelGGL(k_w_zero, zgrid, dim3(kWB), 0, s, d_scr, (const WideState*)d_st);
  hipLaunchKernelGGL(k_w_accum, egrid, dim3(kWB), 0, s, d_scr, (const WideState*)d_st);
  hipLaunchKernelGGL(k_w_rowsum, rgrid, dim3(kWB), 0, s, d_scr, d_st);
  hipLaunchKernelGGL(k_w_bounds, dim3(ctl), dim3(kWCtl), 0, s, d_scr, (const WideState*)d_st, (uint32_t)k, bounds,
                     results);
  CAC_LAUNCH_CHECK(ctx, "wide densify");
  return CAC_OK;
}

}  // namespace cac
