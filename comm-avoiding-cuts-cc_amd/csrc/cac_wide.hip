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
//   1. masked inclusive u64 prefix of the weights;
//   2. D draws (|V|, or the high-concurrency count): Philox(ctr = {trial, round,
//      draw, 0}, key = {seed, 'SAMP'}), r = 1 + floor(x W / 2^64), edge = first
//      prefix >= r;
//   3. union-find over the draws.  When all D draws leave more than t
//      super-vertices they are all kept; otherwise the shortest draw prefix
//      that reaches t is binary-searched.  The search state lives on the
//      device per trial, so a probe step is four launches with no host wait;
//   4. canonical relabel (hooks attach the larger root under the smaller, so
//      every root is its set's minimum id).
// Two things make the prefix cheaper than a full masked scan per round:
//   * round 0's prefix (every label distinct, loops masked) is the same for
//     every trial of a graph: it is built once per graph and searched by all;
//   * from round 1 a trial's scan reads only the edges that survived its
//     previous scan (an edge that became a loop stays one) and writes the new
//     survivors, in order, with the prefix over them.  Dropped edges weigh 0 in
//     the masked prefix, so the first survivor whose prefix reaches r is the
//     edge the full prefix names.
// The host waits twice per round: after the first probe (does any trial
// search?) and after the relabel (which trials go on, and their list lengths).
//
// Union-find here is touched by workgroups on different XCDs, whose L2s are
// not coherent with each other for plain loads within a kernel.  Parent reads
// and path-splitting writes are relaxed agent-scope atomics, so a find always
// sees the hooks other XCDs made and the CAS retry loop makes progress.
#include <algorithm>
#include <unordered_map>
#include <vector>

#include "cac_common.h"
#include "cac_internal.h"
#include "cac_ks.h"
#include "cac_wave.h"
#include "cac_trial.h"

namespace cac {

namespace {

constexpr int kWB = 256;                    // threads per block
constexpr int kWWaves = kWB / 64;
constexpr int kWItems = 16;                 // items per thread in a tile
constexpr uint32_t kWTile = kWB * kWItems;  // edges (or vertices) per tile
constexpr int kWCtl = 64;                   // threads per block of the per-trial control kernels
static_assert(kWItems * kWWaves == 64, "one wave scans a tile's segment totals");

enum : uint32_t { kWIdle = 0, kWKeep = 1, kWSearch = 2, kWFinal = 3, kWDone = 4 };

__host__ __device__ inline uint64_t tiles_of(uint64_t count) { return (count + kWTile - 1) / kWTile; }

// A distinct graph of a wide launch and its round-0 prefix (device).
struct WideGraph {
  const cac_edge* edges;
  uint64_t m;
  uint64_t* pre0;   // [m] inclusive prefix of the weights, loops weigh 0
  uint64_t* tsum0;  // [tiles] inclusive tile totals of pre0
  uint64_t W0;      // total
};

// A wide trial's control words and its scan buffers (device).
struct WideState {
  uint64_t* tsum;     // [tiles] inclusive weight totals of the last scan, by input tile
  uint32_t* tcnt;     // [tiles] inclusive survivor counts of the last scan
  uint32_t* list[2];  // survivor lists (edge indices), written alternately
  uint64_t W;         // the round's masked total
  uint64_t bmin;      // densify: minimum row sum
  uint64_t len;       // survivors of the last scan, in list[out]
  uint64_t len_in;    // entries the last scan read
  uint64_t len_next;  // survivors of the scan in progress
  uint32_t g;         // the trial's graph
  uint32_t out;       // list[out] holds the last scan's survivors
  uint32_t scanned;   // 1 once the trial has scanned its own edges (rounds >= 1)
  uint32_t nv;        // super-vertices
  uint32_t active;    // 1 while the trial is sampling this round
  uint32_t D;         // draws of the round
  uint32_t mode;      // kW* probe state of the round
  uint32_t lo, hi;    // draw-prefix search interval
  uint32_t last;      // length of the last probe (its forest is in par)
  uint32_t cur;       // length of this step's probe (0: none)
  uint32_t undo;      // draws whose endpoints this step resets first
  uint32_t hooks;     // hooks of the last probe
  uint32_t round;
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

__device__ inline uint64_t wave_sum64(uint64_t x) { return wave_sum<uint64_t>(x); }

template <typename T>
__device__ inline T wave_incl(T x, int /*lane*/) {
  return wave_incl_scan<T>(x);
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
  const uint64_t inc = wave_incl(x, lane);
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

// Inclusive scan, in place, of count u64 (and, when c is not null, u32) tile
// totals by one 1024-thread block; the totals go to *tot64 and *totc.
__device__ inline void block_scan_tiles(uint64_t* t64, uint32_t* c, uint64_t count, uint64_t* tot64,
                                        uint64_t* totc) {
  __shared__ uint64_t part[16], partc[16];
  __shared__ uint64_t carry[2];
  if (threadIdx.x == 0) carry[0] = carry[1] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint64_t base = 0; base < count; base += 1024) {
    const uint64_t i = base + threadIdx.x;
    const uint64_t x = i < count ? t64[i] : 0ull;
    const uint64_t y = (c && i < count) ? (uint64_t)c[i] : 0ull;
    const uint64_t inc = wave_incl(x, lane), incc = wave_incl(y, lane);
    if (lane == 63) {
      part[wave] = inc;
      partc[wave] = incc;
    }
    __syncthreads();
    uint64_t off = carry[0], offc = carry[1], tot = 0, totcc = 0;
    for (int v = 0; v < 16; ++v) {
      if (v < wave) {
        off += part[v];
        offc += partc[v];
      }
      tot += part[v];
      totcc += partc[v];
    }
    if (i < count) {
      t64[i] = off + inc;
      if (c) c[i] = (uint32_t)(offc + incc);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      carry[0] += tot;
      carry[1] += totcc;
    }
    __syncthreads();
  }
  *tot64 = carry[0];
  *totc = carry[1];
}

// ---- round 0: the graph's prefix, shared by its trials ---------------------

__device__ inline uint64_t loopless_w(const cac_edge& x) { return x.u != x.v ? x.w : 0ull; }

// Grid (edge tiles, graphs).
__global__ void __launch_bounds__(kWB) k_wg_sums(WideGraph* gs) {
  WideGraph& G = gs[blockIdx.y];
  const uint64_t base = (uint64_t)blockIdx.x * kWTile;
  if (base >= G.m) return;
  uint64_t s = 0;
#pragma unroll 4
  for (int j = 0; j < kWItems; ++j) {
    const uint64_t e = base + (uint64_t)j * kWB + threadIdx.x;
    if (e < G.m) s += loopless_w(G.edges[e]);
  }
  s = block_sum64(s);
  if (threadIdx.x == 0) G.tsum0[blockIdx.x] = s;
}

// Grid (graphs), 1024 threads.
__global__ void __launch_bounds__(1024) k_wg_scan(WideGraph* gs) {
  WideGraph& G = gs[blockIdx.x];
  uint64_t tot = 0, totc = 0;
  block_scan_tiles(G.tsum0, nullptr, tiles_of(G.m), &tot, &totc);
  if (threadIdx.x == 0) G.W0 = tot;
}

// Grid (edge tiles, graphs): pre0 of the tile.  Item j of thread x is edge
// j * kWB + x of the tile, so every wave access covers one contiguous run; in
// tile order the items go by (j, wave, lane), so each wave scans its 64-lane
// segment of every row j and one wave scans the 64 segment totals in order.
__global__ void __launch_bounds__(kWB) k_wg_prefix(WideGraph* gs) {
  WideGraph& G = gs[blockIdx.y];
  const uint64_t base = (uint64_t)blockIdx.x * kWTile;
  if (base >= G.m) return;
  __shared__ uint64_t seg[64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t v[kWItems];
#pragma unroll
  for (int j = 0; j < kWItems; ++j) {
    const uint64_t e = base + (uint64_t)j * kWB + threadIdx.x;
    v[j] = e < G.m ? loopless_w(G.edges[e]) : 0ull;
  }
#pragma unroll
  for (int j = 0; j < kWItems; ++j) {
    v[j] = wave_incl(v[j], lane);
    if (lane == 63) seg[j * kWWaves + wave] = v[j];
  }
  __syncthreads();
  if (wave == 0) {
    const uint64_t x = seg[lane];
    seg[lane] = wave_incl(x, lane) - x + (blockIdx.x ? G.tsum0[blockIdx.x - 1] : 0ull);
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kWItems; ++j) {
    const uint64_t e = base + (uint64_t)j * kWB + threadIdx.x;
    if (e < G.m) G.pre0[e] = seg[j * kWWaves + wave] + v[j];
  }
}

// ---- sampling rounds ----------------------------------------------------------

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
    w.scanned = 0;
    w.out = 0;
    w.len = 0;
    w.cur = 0;
    w.undo = 0;
  }
}

// Per trial, round 0: the graph's total; the first probe takes all D draws.
__global__ void k_w_round0(const TrialScratch* __restrict__ scr, WideState* st, const WideGraph* gs, uint32_t k) {
  const uint32_t i = blockIdx.x * kWCtl + threadIdx.x;
  if (i >= k) return;
  WideState& w = st[i];
  if (!w.active) return;
  w.W = gs[w.g].W0;
  if (w.W == 0) {  // only loops: more than t components
    w.active = 0;
    return;
  }
  w.D = scr[i].dpr ? scr[i].dpr : w.nv;
  w.cur = w.D;
  w.undo = 0;
  w.hooks = 0;
}

// Edge of item i of the trial's scan input: its last survivor list, or every
// edge before its first scan.
__device__ inline uint64_t scan_edge(const WideState& w, uint64_t i) {
  return w.scanned ? (uint64_t)w.list[w.out][i] : i;
}

__device__ inline uint64_t masked_w(const TrialScratch& S, uint64_t e) {
  const cac_edge x = S.edges[e];
  return S.lab[x.u] != S.lab[x.v] ? x.w : 0ull;
}

// Grid (input tiles, trials): per-tile masked totals and survivor counts.
__global__ void __launch_bounds__(kWB) k_w_sums(const TrialScratch* __restrict__ scr, WideState* st) {
  const TrialScratch& S = scr[blockIdx.y];
  WideState& w = st[blockIdx.y];
  if (!w.active) return;
  const uint64_t count = w.scanned ? w.len : S.m;
  const uint64_t base = (uint64_t)blockIdx.x * kWTile;
  if (base >= count) return;
  uint64_t s = 0, c = 0;
#pragma unroll 4
  for (int j = 0; j < kWItems; ++j) {
    const uint64_t i = base + (uint64_t)j * kWB + threadIdx.x;
    if (i < count) {
      const uint64_t x = masked_w(S, scan_edge(w, i));
      s += x;
      c += x ? 1 : 0;
    }
  }
  s = block_sum64(s);
  c = block_sum64(c);
  if (threadIdx.x == 0) {
    w.tsum[blockIdx.x] = s;
    w.tcnt[blockIdx.x] = (uint32_t)c;
  }
}

// Grid (trials), 1024 threads: scans of the tile totals and counts; the
// round's total, survivor count and first probe.  A trial with nothing left to
// draw stops.
__global__ void __launch_bounds__(1024) k_w_scan(const TrialScratch* __restrict__ scr, WideState* st) {
  const TrialScratch& S = scr[blockIdx.x];
  WideState& w = st[blockIdx.x];
  if (!w.active) return;
  const uint64_t count = w.scanned ? w.len : S.m;
  uint64_t tot = 0, totc = 0;
  block_scan_tiles(w.tsum, w.tcnt, tiles_of(count), &tot, &totc);
  if (threadIdx.x == 0) {
    w.len_in = count;
    w.len_next = totc;
    w.W = tot;
    if (tot == 0) {  // only loops left: more than t components
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

// Grid (input tiles, trials): the survivors of the tile, in order, into the
// other list, with the masked inclusive prefix over them.  Item j of thread x
// is input j * kWB + x of the tile (contiguous wave accesses); in tile order
// the items go by (j, wave, lane), so each wave scans its 64-lane segment of
// every row j and one wave scans the 64 segment totals in that order.
__global__ void __launch_bounds__(kWB) k_w_compact(const TrialScratch* __restrict__ scr, const WideState* st) {
  const TrialScratch& S = scr[blockIdx.y];
  const WideState& w = st[blockIdx.y];
  if (!w.active) return;
  const uint64_t count = w.scanned ? w.len : S.m;
  const uint64_t base = (uint64_t)blockIdx.x * kWTile;
  if (base >= count) return;
  __shared__ uint64_t seg[64];
  __shared__ uint32_t segc[64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t v[kWItems];
  uint32_t e[kWItems], c[kWItems];
#pragma unroll
  for (int j = 0; j < kWItems; ++j) {
    const uint64_t i = base + (uint64_t)j * kWB + threadIdx.x;
    e[j] = i < count ? (uint32_t)scan_edge(w, i) : 0u;
    v[j] = i < count ? masked_w(S, e[j]) : 0ull;
  }
#pragma unroll
  for (int j = 0; j < kWItems; ++j) {
    const uint32_t keep = v[j] ? 1u : 0u;
    v[j] = wave_incl(v[j], lane);
    c[j] = wave_incl(keep, lane);
    if (lane == 63) {
      seg[j * kWWaves + wave] = v[j];
      segc[j * kWWaves + wave] = c[j];
    }
    c[j] = keep ? c[j] : 0u;  // 0 marks a dropped item
  }
  __syncthreads();
  if (wave == 0) {
    const uint64_t x = seg[lane];
    const uint32_t y = segc[lane];
    seg[lane] = wave_incl(x, lane) - x + (blockIdx.x ? w.tsum[blockIdx.x - 1] : 0ull);
    segc[lane] = wave_incl(y, lane) - y + (blockIdx.x ? w.tcnt[blockIdx.x - 1] : 0u);
  }
  __syncthreads();
  uint32_t* __restrict__ dst = w.list[w.out ^ 1u];
#pragma unroll
  for (int j = 0; j < kWItems; ++j) {
    if (c[j]) {
      const uint32_t pos = segc[j * kWWaves + wave] + c[j] - 1;
      dst[pos] = e[j];
      S.pre[pos] = seg[j * kWWaves + wave] + v[j];
    }
  }
}

// Per trial, after a scan: its survivors become the current list.
__global__ void k_w_flip(WideState* st, uint32_t k) {
  const uint32_t i = blockIdx.x * kWCtl + threadIdx.x;
  if (i >= k) return;
  WideState& w = st[i];
  if (!w.active) return;
  w.out ^= 1u;
  w.len = w.len_next;
  w.scanned = 1;
}

// Grid (draw blocks, trials): the round's draws.  Before its first scan a
// trial searches its graph's prefix; after, the prefix over its survivors:
// the first input tile whose total reaches r, then that tile's survivors.
__global__ void __launch_bounds__(kWB) k_w_draws(const TrialScratch* __restrict__ scr, const WideState* st,
                                                 const WideGraph* gs) {
  const TrialScratch& S = scr[blockIdx.y];
  const WideState& w = st[blockIdx.y];
  if (!w.active) return;
  const uint32_t d = blockIdx.x * kWB + threadIdx.x;
  if (d >= w.D) return;
  const u32x4 rnd = philox4x32(S.trial, w.round, d, 0u, S.seed, kDomSample);
  const uint64_t r = uniform1(lo64(rnd), w.W);
  const uint64_t* tsum = w.scanned ? w.tsum : gs[w.g].tsum0;
  const uint64_t* pre = w.scanned ? S.pre : gs[w.g].pre0;
  const uint32_t tiles = (uint32_t)tiles_of(w.scanned ? w.len_in : S.m);
  uint32_t a = 0, b = tiles - 1;  // first tile whose inclusive total reaches r
  while (a < b) {
    const uint32_t mid = (a + b) >> 1;
    if (tsum[mid] >= r) b = mid;
    else a = mid + 1;
  }
  uint64_t lo, hi;
  if (w.scanned) {  // tile a keeps at least one survivor: its total grew past r - 1
    lo = a ? w.tcnt[a - 1] : 0u;
    hi = (uint64_t)w.tcnt[a] - 1;
  } else {
    lo = (uint64_t)a * kWTile;
    hi = min(lo + kWTile, S.m) - 1;
  }
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (pre[mid] >= r) hi = mid;
    else lo = mid + 1;
  }
  const uint64_t e = w.scanned ? (uint64_t)w.list[w.out][lo] : lo;
  const cac_edge x = S.edges[e];
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

// Grid (vertex tiles, trials): per-tile counts of canonical labels.
__global__ void __launch_bounds__(kWB) k_w_live_sums(const TrialScratch* __restrict__ scr, WideState* st) {
  const TrialScratch& S = scr[blockIdx.y];
  WideState& w = st[blockIdx.y];
  const uint64_t base = (uint64_t)blockIdx.x * kWTile;
  if (base >= S.n) return;
  uint64_t s = 0;
#pragma unroll 4
  for (int j = 0; j < kWItems; ++j) {
    const uint64_t v = base + (uint64_t)j * kWB + threadIdx.x;
    if (v < S.n) s += S.lab[v] == (uint32_t)v ? 1ull : 0ull;
  }
  s = block_sum64(s);
  if (threadIdx.x == 0) w.tsum[blockIdx.x] = s;
}

// Grid (trials), 1024 threads.
__global__ void __launch_bounds__(1024) k_w_live_scan(const TrialScratch* __restrict__ scr, WideState* st) {
  WideState& w = st[blockIdx.x];
  uint64_t tot = 0, totc = 0;
  block_scan_tiles(w.tsum, nullptr, tiles_of(scr[blockIdx.x].n), &tot, &totc);
  if (threadIdx.x == 0) w.bmin = ~0ull;
}

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

// Grid (input tiles, trials): accumulate the surviving edges into the matrix.
// The last scan's survivors hold every edge that is not a loop now.
__global__ void __launch_bounds__(kWB) k_w_accum(const TrialScratch* __restrict__ scr, const WideState* st) {
  const TrialScratch& S = scr[blockIdx.y];
  const WideState& w = st[blockIdx.y];
  const uint32_t k = w.nv;
  if (k > S.t) return;
  const uint64_t count = w.scanned ? w.len : S.m;
  const uint64_t base = (uint64_t)blockIdx.x * kWTile;
  if (base >= count) return;
#pragma unroll 4
  for (int j = 0; j < kWItems; ++j) {
    const uint64_t i = base + (uint64_t)j * kWB + threadIdx.x;
    if (i >= count) break;
    const cac_edge x = S.edges[scan_edge(w, i)];
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

// Host view of a wide launch's device state, carved from the caller's work
// buffer in one fixed order: the trial states, the graph records, then per
// trial its tile buffers and survivor lists, and after the first trial of each
// graph the graph's round-0 prefix.
struct WideLayout {
  WideState* d_st = nullptr;
  WideGraph* d_gs = nullptr;
  std::vector<WideState> st;
  std::vector<WideGraph> gs;
};

WideLayout wide_layout(const std::vector<TrialScratch>& h, void* work) {
  WideLayout L;
  Carver cv(work);
  L.d_st = cv.take<WideState>(h.size());
  L.d_gs = cv.take<WideGraph>(h.size());
  L.st.assign(h.size(), WideState{});
  std::unordered_map<const cac_edge*, uint32_t> index;
  for (size_t i = 0; i < h.size(); ++i) {
    WideState& w = L.st[i];
    const uint64_t tiles = tiles_of(std::max<uint64_t>(h[i].m, h[i].n)) + 1;
    w.tsum = cv.take<uint64_t>(tiles);
    w.tcnt = cv.take<uint32_t>(tiles);
    w.list[0] = cv.take<uint32_t>(std::max<uint64_t>(h[i].m, 1));
    w.list[1] = cv.take<uint32_t>(std::max<uint64_t>(h[i].m, 1));
    auto it = index.find(h[i].edges);
    if (it == index.end()) {
      WideGraph g{};
      g.edges = h[i].edges;
      g.m = h[i].m;
      g.pre0 = cv.take<uint64_t>(std::max<uint64_t>(h[i].m, 1));
      g.tsum0 = cv.take<uint64_t>(tiles_of(h[i].m) + 1);
      it = index.emplace(h[i].edges, (uint32_t)L.gs.size()).first;
      L.gs.push_back(g);
    }
    w.g = it->second;
  }
  return L;
}

}  // namespace

uint64_t wide_min_edges(const cac_ctx* ctx) {
  const int32_t v = ctx->wide_log2;
  if (v == 0) return 1ull << 20;
  if (v >= 63) return ~0ull;
  return 1ull << v;
}

size_t wide_trial_bytes(uint64_t m, uint32_t n) {
  const uint64_t tiles = tiles_of(std::max<uint64_t>(m, n)) + 1;
  const uint64_t m1 = std::max<uint64_t>(m, 1);
  return carve_size(sizeof(WideState)) + carve_size(sizeof(WideGraph)) + carve_size(tiles * 8) +
         carve_size(tiles * 4) + 2 * carve_size(m1 * 4) + carve_size(m1 * 8) + carve_size((tiles_of(m) + 1) * 8);
}

int wide_sample(cac_ctx* ctx, const std::vector<TrialScratch>& h, const TrialScratch* d_scr, void* work,
                hipStream_t s) {
  const size_t k = h.size();
  if (k == 0) return CAC_OK;
  if (k > 65535) return fail(ctx, CAC_EINVAL, "too many wide trials in one launch");
  for (const TrialScratch& x : h)
    if (x.m >= (1ull << 32)) return fail(ctx, CAC_EINVAL, "wide trials take fewer than 2^32 edges");
  WideLayout L = wide_layout(h, work);
  uint64_t mmax = 1, nmax = 1;
  for (const TrialScratch& x : h) {
    mmax = std::max<uint64_t>(mmax, x.m);
    nmax = std::max<uint64_t>(nmax, x.n);
  }
  const size_t ng = L.gs.size();
  CAC_HIP(ctx, hipMemcpyAsync(L.d_st, L.st.data(), k * sizeof(WideState), hipMemcpyHostToDevice, s));
  CAC_HIP(ctx, hipMemcpyAsync(L.d_gs, L.gs.data(), ng * sizeof(WideGraph), hipMemcpyHostToDevice, s));
  const unsigned ctl = (unsigned)((k + kWCtl - 1) / kWCtl);
  const dim3 vgrid((unsigned)((nmax + kWB - 1) / kWB), (unsigned)k);
  const dim3 ggrid((unsigned)tiles_of(mmax), (unsigned)ng);
  hipLaunchKernelGGL(k_w_init, vgrid, dim3(kWB), 0, s, d_scr, L.d_st);
  hipLaunchKernelGGL(k_wg_sums, ggrid, dim3(kWB), 0, s, L.d_gs);
  hipLaunchKernelGGL(k_wg_scan, dim3((unsigned)ng), dim3(1024), 0, s, L.d_gs);
  hipLaunchKernelGGL(k_wg_prefix, ggrid, dim3(kWB), 0, s, L.d_gs);
  CAC_LAUNCH_CHECK(ctx, "wide sampling init");
  std::vector<WideState> now(k);
  CAC_HIP(ctx, hipMemcpyAsync(now.data(), L.d_st, k * sizeof(WideState), hipMemcpyDeviceToHost, s));
  CAC_HIP(ctx, hipStreamSynchronize(s));
  for (uint32_t round = 0;; ++round) {
    uint64_t dmax = 0, lmax = 0;
    for (size_t i = 0; i < k; ++i) {
      if (!now[i].active) continue;
      dmax = std::max<uint64_t>(dmax, h[i].dpr ? h[i].dpr : now[i].nv);
      lmax = std::max<uint64_t>(lmax, now[i].scanned ? now[i].len : h[i].m);
    }
    if (dmax == 0) break;
    const dim3 dgrid((unsigned)((dmax + kWB - 1) / kWB), (unsigned)k);
    if (round == 0) {
      hipLaunchKernelGGL(k_w_round0, dim3(ctl), dim3(kWCtl), 0, s, d_scr, L.d_st, (const WideGraph*)L.d_gs,
                         (uint32_t)k);
    } else {
      const dim3 egrid((unsigned)std::max<uint64_t>(tiles_of(lmax), 1), (unsigned)k);
      hipLaunchKernelGGL(k_w_sums, egrid, dim3(kWB), 0, s, d_scr, L.d_st);
      hipLaunchKernelGGL(k_w_scan, dim3((unsigned)k), dim3(1024), 0, s, d_scr, L.d_st);
      hipLaunchKernelGGL(k_w_compact, egrid, dim3(kWB), 0, s, d_scr, (const WideState*)L.d_st);
      hipLaunchKernelGGL(k_w_flip, dim3(ctl), dim3(kWCtl), 0, s, L.d_st, (uint32_t)k);
    }
    hipLaunchKernelGGL(k_w_draws, dgrid, dim3(kWB), 0, s, d_scr, (const WideState*)L.d_st, (const WideGraph*)L.d_gs);
    hipLaunchKernelGGL(k_w_union, dgrid, dim3(kWB), 0, s, d_scr, L.d_st);
    hipLaunchKernelGGL(k_w_after_first, dim3(ctl), dim3(kWCtl), 0, s, d_scr, L.d_st, (uint32_t)k);
    CAC_LAUNCH_CHECK(ctx, "wide sampling round");
    CAC_HIP(ctx, hipMemcpyAsync(now.data(), L.d_st, k * sizeof(WideState), hipMemcpyDeviceToHost, s));
    CAC_HIP(ctx, hipStreamSynchronize(s));
    bool search = false;
    for (size_t i = 0; i < k; ++i) search = search || (now[i].active && now[i].mode == kWSearch);
    if (search) {
      // a search over [1, D] ends within ceil(log2 D) probes, plus the final one
      int steps = 2;
      while ((1ull << (steps - 2)) < dmax) ++steps;
      for (int it = 0; it < steps; ++it) {
        hipLaunchKernelGGL(k_w_step_begin, dim3(ctl), dim3(kWCtl), 0, s, L.d_st, (uint32_t)k);
        hipLaunchKernelGGL(k_w_undo, dgrid, dim3(kWB), 0, s, d_scr, (const WideState*)L.d_st);
        hipLaunchKernelGGL(k_w_union, dgrid, dim3(kWB), 0, s, d_scr, L.d_st);
        hipLaunchKernelGGL(k_w_step_end, dim3(ctl), dim3(kWCtl), 0, s, d_scr, L.d_st, (uint32_t)k);
      }
      CAC_LAUNCH_CHECK(ctx, "wide sampling search");
    }
    hipLaunchKernelGGL(k_w_relabel, vgrid, dim3(kWB), 0, s, d_scr, (const WideState*)L.d_st);
    hipLaunchKernelGGL(k_w_round_end, dim3(ctl), dim3(kWCtl), 0, s, d_scr, L.d_st, (uint32_t)k);
    CAC_LAUNCH_CHECK(ctx, "wide sampling relabel");
    CAC_HIP(ctx, hipMemcpyAsync(now.data(), L.d_st, k * sizeof(WideState), hipMemcpyDeviceToHost, s));
    CAC_HIP(ctx, hipStreamSynchronize(s));
    for (size_t i = 0; i < k; ++i)
      if (now[i].mode != kWIdle) return fail(ctx, CAC_EDEVICE, "wide sampling: draw-prefix search did not end");
  }
  hipLaunchKernelGGL(k_w_nv_out, dim3(ctl), dim3(kWCtl), 0, s, d_scr, (const WideState*)L.d_st, (uint32_t)k);
  CAC_LAUNCH_CHECK(ctx, "wide sampling result");
  return CAC_OK;
}

int wide_densify(cac_ctx* ctx, const std::vector<TrialScratch>& h, const TrialScratch* d_scr, void* work,
                 int64_t* bounds, int64_t* results, hipStream_t s) {
  const size_t k = h.size();
  if (k == 0) return CAC_OK;
  WideLayout L = wide_layout(h, work);
  std::vector<WideState> now(k);
  CAC_HIP(ctx, hipMemcpyAsync(now.data(), L.d_st, k * sizeof(WideState), hipMemcpyDeviceToHost, s));
  CAC_HIP(ctx, hipStreamSynchronize(s));
  uint64_t lmax = 1, nmax = 1, kmax = 1;
  for (size_t i = 0; i < k; ++i) {
    lmax = std::max<uint64_t>(lmax, now[i].scanned ? now[i].len : h[i].m);
    nmax = std::max<uint64_t>(nmax, h[i].n);
    if (now[i].nv <= h[i].t) kmax = std::max<uint64_t>(kmax, now[i].nv);
  }
  const unsigned ctl = (unsigned)((k + kWCtl - 1) / kWCtl);
  const dim3 tgrid((unsigned)tiles_of(nmax), (unsigned)k);
  const dim3 egrid((unsigned)tiles_of(lmax), (unsigned)k);
  const uint64_t cells = kmax * kmax;
  const dim3 zgrid((unsigned)std::min<uint64_t>((cells + kWB * 8 - 1) / (kWB * 8), 65536), (unsigned)k);
  const dim3 rgrid((unsigned)((kmax + kWWaves - 1) / kWWaves), (unsigned)k);
  hipLaunchKernelGGL(k_w_live_sums, tgrid, dim3(kWB), 0, s, d_scr, L.d_st);
  hipLaunchKernelGGL(k_w_live_scan, dim3((unsigned)k), dim3(1024), 0, s, d_scr, L.d_st);
  hipLaunchKernelGGL(k_w_rank, tgrid, dim3(kWB), 0, s, d_scr, (const WideState*)L.d_st);
  hipLaunchKernelGGL(k_w_zero, zgrid, dim3(kWB), 0, s, d_scr, (const WideState*)L.d_st);
  hipLaunchKernelGGL(k_w_accum, egrid, dim3(kWB), 0, s, d_scr, (const WideState*)L.d_st);
  hipLaunchKernelGGL(k_w_rowsum, rgrid, dim3(kWB), 0, s, d_scr, L.d_st);
  hipLaunchKernelGGL(k_w_bounds, dim3(ctl), dim3(kWCtl), 0, s, d_scr, (const WideState*)L.d_st, (uint32_t)k, bounds,
                     results);
  CAC_LAUNCH_CHECK(ctx, "wide densify");
  return CAC_OK;
}

}  // namespace cac
