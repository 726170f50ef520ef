// cac_sparse.hip -- sparse Karger-Stein on gfx950 and the karger_stein driver.
//
// Replaces the sparse recursion of comincut: contract_graph / sparse_contract_edges
// (src/karger-stein/co_mincut.h:74-155), recursive_contraction on sparse graphs
// (:176-223, :280-293), the dense/sparse dispatch of minimum_cut_trials_start
// (:396-414) and SequentialKargerSteinCut::compute (src/SequentialKargerSteinCut.cpp:6-27).
//
// A node is an edge list (u < v, w >= 1, parallel edges allowed) over n vertices.
// Contracting it to t' vertices is an exponential race: edge e gets the integer
// key race_key(x_e, w_e), x_e from Philox(ctr = {trial, node id, e_lo, e_hi},
// key = {seed, 'SPRS'}), and edges are united in (key, index) order until t'
// components remain.  One 1024-thread workgroup per child finds that prefix by
// narrowing a candidate key range, as the reference halves around nth_element:
// the first cut aims at the quantile that should hold twice the unions still
// needed, later cuts at the median, and a trial union-find kept equal to the
// identity between probes makes each probe cost only its own edges.  The child
// is the node's edges in their order, relabelled by the rank of each
// component's minimum id, loops dropped and parallel edges kept.  A node with
// n <= 2048, or with m >= n^2/32 and n <= 8192, is densified (band by band in
// LDS: every matrix entry written once) and continues in the dense engine under
// its own node id.  Every node uses the trial's bound,
// the minimum weighted degree of the input.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "cac_common.h"
#include "cac_internal.h"
#include "cac_ks.h"
#include "cac_ks_kernels.h"

namespace cac {
namespace {

constexpr uint32_t kDomSparse = 0x53505253u;  // 'SPRS'
constexpr int kSpBlock = 1024;
constexpr int kSpWaves = kSpBlock / 64;
constexpr uint32_t kSpSort = 1024;  // LDS sort size: pivot samples, final candidates
constexpr int kSpRows = 4;          // a compaction tile is kSpRows x kSpBlock elements, row after row
constexpr uint64_t kSpTile = (uint64_t)kSpRows * kSpBlock;  // edges per output tile
constexpr uint32_t kSparseDenseMax = 2048;
constexpr unsigned long long kSpDisconnected = ~0ull;

}  // namespace

// floor(2^20 * L / w), L = -log2(U / 2^53) in Q32.32, U = (x >> 11) + 1.  The
// fraction of log2 comes from bit-by-bit squaring of the top 32 mantissa bits
// (each square truncated to 32 bits, |error| < 2^-29), so host and device agree
// exactly.
__host__ __device__ inline uint64_t race_key(uint64_t x, uint64_t w) {
  const uint64_t U = (x >> 11) + 1;
#if defined(__HIP_DEVICE_COMPILE__)
  const int e = 63 - __clzll((long long)U);
#else
  const int e = 63 - __builtin_clzll(U);
#endif
  uint32_t y = (uint32_t)((U << (63 - e)) >> 32);  // mantissa in [1, 2), Q1.31
  uint32_t frac = 0;
#pragma unroll 8
  for (int i = 0; i < 32; ++i) {
    const uint64_t sq = (uint64_t)y * y;  // Q2.62
    const bool top = (sq >> 63) != 0;
    frac = (frac << 1) | (top ? 1u : 0u);
    y = top ? (uint32_t)(sq >> 32) : (uint32_t)(sq >> 31);
  }
  const uint64_t L = ((uint64_t)53 << 32) - (((uint64_t)e << 32) | frac);
  return w == 1 ? (L << 20) : (L << 20) / (w ? w : 1);
}

namespace {

struct SpChild {
  const cac_edge* edges;  // parent edges [m], u < v
  uint64_t m;
  uint32_t n, tgt, trial, nid;  // nid: this child's node id (keys come from it)
  uint32_t slot, seed;          // seed: Philox key word 0 of the child's trial
  cac_edge* out;              // [m] child edges (kept for the next depth)
  uint64_t* keys;             // [m] scratch: race keys
  uint32_t* list[2];          // [m] scratch: candidate edge indices, ping-pong
  uint32_t* parc;             // [n] scratch: committed forest, then component ranks
  uint32_t* part;             // [n] scratch: trial forest (identity between probes), then labels
  unsigned long long* m_out;  // child edge count, kSpDisconnected when disconnected
  unsigned long long* kmm;    // [2] smallest and largest race key, from k_sp_keys
  uint32_t* tiles;            // [ceil(m / kSpTile)] scratch: kept edges per output tile, then their offsets
};

// Root with path halving.  Links always point to smaller ids, so even a stale
// view of the forest ends at a root.
__device__ inline uint32_t sp_root(uint32_t* par, uint32_t x) {
  uint32_t p = par[x];
  while (p != x) {
    const uint32_t g = par[p];
    if (g != p) par[x] = g;
    x = p;
    p = g;
  }
  return x;
}

// Links the sets of a and b (larger root under the smaller); true when this
// call joined two sets.
__device__ inline bool sp_union(uint32_t* par, uint32_t a, uint32_t b) {
  uint32_t ra = sp_root(par, a), rb = sp_root(par, b);
  while (ra != rb) {
    const uint32_t lo = min(ra, rb), hi = max(ra, rb);
    const uint32_t old = atomicCAS(par + hi, hi, lo);
    if (old == hi) return true;
    ra = sp_root(par, old);
    rb = sp_root(par, lo);
  }
  return false;
}

__device__ inline uint64_t sp_sum(uint64_t x, uint64_t* red) { return block_sum_u64<kSpBlock>(x, red); }

__device__ inline uint64_t sp_min(uint64_t x, uint64_t* red) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) x = min(x, (uint64_t)__shfl_xor(x, d, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  uint64_t t = ~0ull;
  for (int w = 0; w < kSpWaves; ++w) t = min(t, red[w]);
  __syncthreads();
  return t;
}

__device__ inline uint64_t sp_max(uint64_t x, uint64_t* red) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) x = max(x, (uint64_t)__shfl_xor(x, d, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  uint64_t t = 0;
  for (int w = 0; w < kSpWaves; ++w) t = max(t, red[w]);
  __syncthreads();
  return t;
}

// Exclusive scan of per-thread counts in thread order; *total gets the sum.
__device__ inline uint64_t sp_exscan(uint64_t x, uint64_t* red, uint64_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t inc = wave_incl_scan<uint64_t>(x);
  if (lane == 63) red[wave] = inc;
  __syncthreads();
  uint64_t before = 0, t = 0;
  for (int w = 0; w < kSpWaves; ++w) {
    if (w < wave) before += red[w];
    t += red[w];
  }
  __syncthreads();
  *total = t;
  return before + inc - x;
}

// Stable compaction ranks of a tile of kSpRows rows of kSpBlock elements (row j
// of the thread is element j * kSpBlock + threadIdx.x of the tile): the
// kSpRows x kSpWaves wave counts, in tile order, are one wave's exclusive scan,
// so a tile of kSpRows rows costs three barriers instead of two per row.
// wcnt holds kSpRows * kSpWaves + 1 words.
static_assert(kSpRows * kSpWaves == 64, "one scan lane per (row, wave) count");
__device__ inline void sp_tile_ranks(const bool (&keep)[kSpRows], uint32_t (&rank)[kSpRows], uint32_t* wcnt,
                                     uint32_t* tile_total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t below = (1ull << lane) - 1;
#pragma unroll
  for (int j = 0; j < kSpRows; ++j) {
    const uint64_t bal = __ballot(keep[j]);
    rank[j] = (uint32_t)__popcll(bal & below);
    if (lane == 0) wcnt[j * kSpWaves + wave] = (uint32_t)__popcll(bal);
  }
  __syncthreads();
  if (wave == 0) {
    const uint32_t v = wcnt[lane];
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(inc, d, 64);
      if (lane >= d) inc += y;
    }
    wcnt[lane] = inc - v;
    if (lane == 63) wcnt[64] = inc;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSpRows; ++j) rank[j] += wcnt[j * kSpWaves + wave];
  *tile_total = wcnt[64];
  __syncthreads();  // the next tile rewrites wcnt
}

// Bitonic sort of kSpSort (key, idx) pairs in LDS, ascending by key then idx.
__device__ inline void sp_sort(uint64_t* key, uint32_t* idx) {
  const uint32_t i = threadIdx.x;
  for (uint32_t k = 2; k <= kSpSort; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      const uint32_t p = i ^ j;
      if (p > i) {
        const uint64_t ka = key[i], kb = key[p];
        const uint32_t ia = idx[i], ib = idx[p];
        const bool greater = ka > kb || (ka == kb && ia > ib);
        const bool up = (i & k) == 0;
        if (greater == up) {
          key[i] = kb;
          key[p] = ka;
          idx[i] = ib;
          idx[p] = ia;
        }
      }
      __syncthreads();
    }
  }
}

// Candidate edge at position pos: list[pos], or pos itself while the candidates
// are still every edge (list == nullptr).
__device__ inline uint32_t sp_at(const uint32_t* list, uint64_t pos) { return list ? list[pos] : (uint32_t)pos; }

// Stable compaction of the candidates whose key satisfies keep into dst; returns
// the count, and the smallest and largest kept key in *kmin, *kmax.
template <class Keep>
__device__ uint64_t sp_filter(const SpChild& c, const uint32_t* src, uint64_t ns, Keep keep, uint32_t* dst,
                              uint64_t* kmin, uint64_t* kmax, uint32_t* wcnt, uint64_t* red) {
  uint64_t base = 0, lo = ~0ull, hi = 0;
  for (uint64_t t0 = 0; t0 < ns; t0 += (uint64_t)kSpBlock * kSpRows) {
    uint32_t e[kSpRows], r[kSpRows];
    bool k[kSpRows];
#pragma unroll
    for (int j = 0; j < kSpRows; ++j) {
      const uint64_t pos = t0 + (uint64_t)j * kSpBlock + threadIdx.x;
      e[j] = 0;
      k[j] = false;
      if (pos < ns) {
        e[j] = sp_at(src, pos);
        const uint64_t key = c.keys[e[j]];
        k[j] = keep(key);
        if (k[j]) {
          lo = min(lo, key);
          hi = max(hi, key);
        }
      }
    }
    uint32_t tot = 0;
    sp_tile_ranks(k, r, wcnt, &tot);
#pragma unroll
    for (int j = 0; j < kSpRows; ++j)
      if (k[j]) dst[base + r[j]] = e[j];
    base += tot;
  }
  *kmin = sp_min(lo, red);
  *kmax = sp_max(hi, red);
  block_sync_global();
  return base;
}

// Components removed by uniting, on top of the committed forest, the first p
// candidates.  The trial forest is restored to the identity afterwards: every
// entry a probe writes is the committed root of an endpoint of one of its
// edges, and those are reset.
__device__ uint32_t sp_probe(const SpChild& c, const uint32_t* list, uint64_t p, uint64_t* red) {
  uint64_t merged = 0;
  for (uint64_t pos = threadIdx.x; pos < p; pos += kSpBlock) {
    const cac_edge x = c.edges[sp_at(list, pos)];
    const uint32_t a = sp_root(c.parc, x.u), b = sp_root(c.parc, x.v);
    if (a != b && sp_union(c.part, a, b)) ++merged;
  }
  block_sync_global();
  for (uint64_t pos = threadIdx.x; pos < p; pos += kSpBlock) {
    const cac_edge x = c.edges[sp_at(list, pos)];
    const uint32_t a = sp_root(c.parc, x.u), b = sp_root(c.parc, x.v);
    c.part[a] = a;
    c.part[b] = b;
  }
  const uint64_t total = sp_sum(merged, red);
  block_sync_global();
  return (uint32_t)total;
}

// Unites the first p candidates into the committed forest.
__device__ void sp_commit(const SpChild& c, const uint32_t* list, uint64_t p) {
  for (uint64_t pos = threadIdx.x; pos < p; pos += kSpBlock) {
    const cac_edge x = c.edges[sp_at(list, pos)];
    sp_union(c.parc, x.u, x.v);
  }
  block_sync_global();
}

struct KeepBelow {  // key <= tau
  uint64_t tau;
  __device__ bool operator()(uint64_t k) const { return k <= tau; }
};
struct KeepAbove {  // key > tau
  uint64_t tau;
  __device__ bool operator()(uint64_t k) const { return k > tau; }
};

__global__ void k_sp_kmm_init(const SpChild* __restrict__ children, uint32_t count) {
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < count; i += gridDim.x * 256) {
    __hip_atomic_store(&children[i].kmm[0], ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&children[i].kmm[1], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(children[i].m_out, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Race keys of every child's parent edges, spread over the whole device (grid
// y = child), with each child's smallest and largest key.
__global__ void __launch_bounds__(256) k_sp_keys(const SpChild* __restrict__ children) {
  const SpChild c = children[blockIdx.y];
  uint64_t kmin = ~0ull, kmax = 0;
  for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < c.m; e += (uint64_t)gridDim.x * 256) {
    const u32x4 r = philox4x32(c.trial, c.nid, (uint32_t)e, (uint32_t)(e >> 32), c.seed, kDomSparse);
    const uint64_t k = race_key(lo64(r), c.edges[e].w);
    c.keys[e] = k;
    kmin = min(kmin, k);
    kmax = max(kmax, k);
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    kmin = min(kmin, (uint64_t)__shfl_xor(kmin, d, 64));
    kmax = max(kmax, (uint64_t)__shfl_xor(kmax, d, 64));
  }
  if ((threadIdx.x & 63) == 0 && kmin <= kmax) {
    atomicMin(&c.kmm[0], (unsigned long long)kmin);
    atomicMax(&c.kmm[1], (unsigned long long)kmax);
  }
}

__global__ void __launch_bounds__(kSpBlock, 8) k_sparse_contract(const SpChild* __restrict__ children,
                                                              int64_t* __restrict__ results) {
  __shared__ uint64_t red[kSpWaves];
  __shared__ uint32_t wcnt[kSpRows * kSpWaves + 1];
  __shared__ uint64_t skey[kSpSort];
  __shared__ uint32_t sidx[kSpSort];
  const SpChild c = children[blockIdx.x];
  const uint32_t tid = threadIdx.x;

  // 1. identity forests; the race keys and their range come from k_sp_keys
  for (uint32_t v = tid; v < c.n; v += kSpBlock) {
    c.parc[v] = v;
    c.part[v] = v;
  }
  uint64_t kmin = __hip_atomic_load(&c.kmm[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint64_t kmax = __hip_atomic_load(&c.kmm[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  block_sync_global();

  if (c.n > c.tgt) {  // else nothing is contracted
    // 2. narrowing: the committed forest holds the edges with key <= lo, the
    //    candidates are the edges with key in (lo, hi]; components > tgt always.
    //    cur == nullptr: the candidates are every edge, in index order.
    const uint32_t* cur = nullptr;
    int nxt_i = 0;
    uint64_t ns = c.m;
    uint32_t comps = c.n;
    uint64_t rho = 2;  // the next cut aims at rho x (unions still needed) edges
    while (ns > kSpSort && kmin < kmax) {
      skey[tid] = c.keys[sp_at(cur, (uint64_t)tid * ns / kSpSort)];
      sidx[tid] = tid;
      __syncthreads();
      sp_sort(skey, sidx);
      const uint64_t want = (rho * (comps - c.tgt) * kSpSort + ns - 1) / ns;  // samples below the cut
      const uint32_t k = (uint32_t)max<uint64_t>(1, min<uint64_t>(want, kSpSort / 2)) - 1;
      const uint64_t tau = min(max(skey[k], kmin), kmax - 1);
      __syncthreads();
      uint32_t* nxt = nxt_i ? c.list[1] : c.list[0];
      uint64_t xmin, xmax;
      const uint64_t nx = sp_filter(c, cur, ns, KeepBelow{tau}, nxt, &xmin, &xmax, wcnt, red);
      const uint32_t merged = sp_probe(c, nxt, nx, red);
      if (comps - merged > c.tgt) {  // not enough: keep these unions, continue above tau
        sp_commit(c, nxt, nx);
        comps -= merged;
        uint64_t rmin, rmax;
        ns = sp_filter(c, cur, ns, KeepAbove{tau}, nxt, &rmin, &rmax, wcnt, red);
        kmin = rmin;
        kmax = rmax;
        rho = min<uint64_t>(rho * 2, 1u << 20);
      } else {
        ns = nx;
        kmin = xmin;
        kmax = xmax;
      }
      cur = nxt;
      nxt_i ^= 1;
    }

    // 3. candidates in (key, index) order: sorted in LDS when few, else they all
    //    share one key and are already in index order
    if (ns <= kSpSort) {
      const bool have = tid < ns;
      const uint32_t e = have ? sp_at(cur, tid) : 0xFFFFFFFFu;
      skey[tid] = have ? c.keys[e] : ~0ull;
      sidx[tid] = e;
      __syncthreads();
      sp_sort(skey, sidx);
      uint32_t* dst = nxt_i ? c.list[1] : c.list[0];
      if (have) dst[tid] = sidx[tid];
      cur = dst;
      block_sync_global();
    }
    // shortest prefix of the candidates that leaves tgt components
    if (comps - sp_probe(c, cur, ns, red) > c.tgt) {
      if (tid == 0) {
        *c.m_out = kSpDisconnected;
        atomicMin((long long*)&results[c.slot], 0ll);
      }
      return;
    }
    uint64_t lo = 0, hi = ns;  // comps(lo) > tgt >= comps(hi)
    while (hi - lo > 1) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if (comps - sp_probe(c, cur, mid, red) <= c.tgt) hi = mid;
      else lo = mid;
    }
    sp_commit(c, cur, hi);
  }

  // 4. labels: part[v] = root, parc[root] = rank of the root (minimum id) in id
  //    order, then part[v] = rank of v's root
  for (uint32_t v = tid; v < c.n; v += kSpBlock) c.part[v] = sp_root(c.parc, v);
  block_sync_global();
  const uint64_t vc = ((uint64_t)c.n + kSpBlock - 1) / kSpBlock;
  const uint32_t v0 = (uint32_t)min((uint64_t)c.n, tid * vc), v1 = (uint32_t)min((uint64_t)c.n, v0 + vc);
  uint64_t cnt = 0;
  for (uint32_t v = v0; v < v1; ++v) cnt += c.part[v] == v;
  uint64_t total = 0;
  uint64_t off = sp_exscan(cnt, red, &total);
  for (uint32_t v = v0; v < v1; ++v)
    if (c.part[v] == v) c.parc[v] = (uint32_t)off++;
  block_sync_global();
  for (uint32_t v = tid; v < c.n; v += kSpBlock) c.part[v] = c.parc[c.part[v]];
  block_sync_global();

  // 5. the child's edges are written by k_sp_tile_count / k_sp_tile_scan /
  //    k_sp_tile_write over the whole device
}

// 5. The child: the parent's edges in parent order, relabelled by the
// component ranks in part, loops dropped; in three launches over the whole
// device (tile counts, their offsets, the stable write).  A disconnected child
// (m_out already kSpDisconnected) is skipped.
__device__ inline bool sp_disconnected(const SpChild& c) {
  return __hip_atomic_load(c.m_out, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == kSpDisconnected;
}

// Grid (tile, child), 256 threads: kept edges of each tile.
__global__ void __launch_bounds__(256) k_sp_tile_count(const SpChild* __restrict__ children) {
  __shared__ uint32_t wsum[4];
  const SpChild c = children[blockIdx.y];
  const uint64_t t0 = (uint64_t)blockIdx.x * kSpTile;
  if (t0 >= c.m || sp_disconnected(c)) return;
  const uint64_t t1 = min(c.m, t0 + kSpTile);
  uint32_t cnt = 0;
  for (uint64_t e = t0 + threadIdx.x; e < t1; e += 256) {
    const cac_edge x = c.edges[e];
    cnt += c.part[x.u] != c.part[x.v] ? 1u : 0u;
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) cnt += __shfl_xor(cnt, d, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) c.tiles[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// One workgroup per child: tile counts to exclusive offsets, and the child's
// edge count.
__global__ void __launch_bounds__(kSpBlock) k_sp_tile_scan(const SpChild* __restrict__ children) {
  __shared__ uint64_t red[kSpWaves];
  const SpChild c = children[blockIdx.x];
  if (sp_disconnected(c)) return;
  const uint64_t nt = (c.m + kSpTile - 1) / kSpTile;
  uint64_t carry = 0;
  for (uint64_t t0 = 0; t0 < nt; t0 += kSpBlock) {
    const uint64_t t = t0 + threadIdx.x;
    const uint64_t v = t < nt ? c.tiles[t] : 0;
    uint64_t total = 0;
    const uint64_t ex = sp_exscan(v, red, &total);
    if (t < nt) c.tiles[t] = (uint32_t)(carry + ex);
    carry += total;
  }
  if (threadIdx.x == 0) *c.m_out = carry;
}

// Grid (tile, child), kSpBlock threads: each tile's kept edges at its offset,
// in order.
__global__ void __launch_bounds__(kSpBlock) k_sp_tile_write(const SpChild* __restrict__ children) {
  __shared__ uint32_t wcnt[kSpRows * kSpWaves + 1];
  const SpChild c = children[blockIdx.y];
  const uint64_t t0 = (uint64_t)blockIdx.x * kSpTile;
  if (t0 >= c.m || sp_disconnected(c)) return;
  const uint64_t base = c.tiles[blockIdx.x];
  uint32_t a[kSpRows], b[kSpRows], r[kSpRows];
  bool keep[kSpRows];
#pragma unroll
  for (int j = 0; j < kSpRows; ++j) {
    const uint64_t e = t0 + (uint64_t)j * kSpBlock + threadIdx.x;
    keep[j] = false;
    a[j] = b[j] = 0;
    if (e < c.m) {
      const cac_edge x = c.edges[e];
      const uint32_t pa = c.part[x.u], pb = c.part[x.v];
      keep[j] = pa != pb;
      a[j] = min(pa, pb);
      b[j] = max(pa, pb);
    }
  }
  uint32_t tot = 0;
  sp_tile_ranks(keep, r, wcnt, &tot);
#pragma unroll
  for (int j = 0; j < kSpRows; ++j)
    if (keep[j]) c.out[base + r[j]] = cac_edge{a[j], b[j], c.edges[t0 + (uint64_t)j * kSpBlock + threadIdx.x].w};
}

// The contraction of nchild children (device descriptors d_children, largest
// parent edge count mmax): keys over the whole device, then one workgroup per
// child.
void launch_sparse_contract(const SpChild* d_children, uint32_t nchild, uint64_t mmax, int64_t* d_results,
                            hipStream_t s) {
  const uint64_t per_child = std::max<uint64_t>(1, 8192 / std::max<uint32_t>(nchild, 1));
  const unsigned gx = (unsigned)std::min<uint64_t>(std::max<uint64_t>(1, (mmax + 4095) / 4096), per_child);
  hipLaunchKernelGGL(k_sp_kmm_init, dim3(grid_for(nchild, 256, 64)), dim3(256), 0, s, d_children, nchild);
  for (uint32_t c0 = 0; c0 < nchild; c0 += 65535u)  // grid y holds at most 65535 children
    hipLaunchKernelGGL(k_sp_keys, dim3(gx, std::min(nchild - c0, 65535u)), dim3(256), 0, s, d_children + c0);
  hipLaunchKernelGGL(k_sparse_contract, dim3(nchild), dim3(kSpBlock), 0, s, d_children, d_results);
  const unsigned tiles = (unsigned)std::max<uint64_t>(1, (mmax + kSpTile - 1) / kSpTile);
  for (uint32_t c0 = 0; c0 < nchild; c0 += 65535u)
    hipLaunchKernelGGL(k_sp_tile_count, dim3(tiles, std::min(nchild - c0, 65535u)), dim3(256), 0, s, d_children + c0);
  hipLaunchKernelGGL(k_sp_tile_scan, dim3(nchild), dim3(kSpBlock), 0, s, d_children);
  for (uint32_t c0 = 0; c0 < nchild; c0 += 65535u)
    hipLaunchKernelGGL(k_sp_tile_write, dim3(tiles, std::min(nchild - c0, 65535u)), dim3(kSpBlock), 0, s,
                       d_children + c0);
}

// Handover to the dense engine: n x n matrix and row sums from an edge list,
// built band by band.  A band is `rows` consecutive rows (rows n <= kSpBandEntries,
// so a band's entries fit LDS).  The edge endpoints are bucketed by band (count,
// exclusive scan, place: entry weight and its index inside the band), then a
// workgroup sums each band's entries in LDS and writes the band's rows, zeros
// included, and their row sums, once.  Every matrix entry is written once and
// never read back; the only atomics are on LDS and on per-tile band cursors.
constexpr uint32_t kSpBandEntries = 8192;                       // int64 entries of a band (64 KiB of LDS)
constexpr int kSpBucketBlock = 1024;                            // bucketing: edges per thread x block = a tile
constexpr int kSpBucketEdges = 8;
constexpr uint64_t kSpBucketTile = (uint64_t)kSpBucketBlock * kSpBucketEdges;
constexpr int kSpBuildBlock = 512;
constexpr size_t kSpBucketLdsMax = (size_t)kKsMaxS * 12;        // bands <= n <= kKsMaxS: u64 base + u32 count

struct SpDense {
  const cac_edge* edges;
  uint64_t m;
  uint32_t n, rows;              // rows per band
  uint32_t bands, pad;           // ceil(n / rows)
  int64_t* mat;                  // n x n
  int64_t* rsum;                 // [n]
  unsigned long long* boff;      // [bands + 1]: entry counts, then their exclusive prefix (zeroed first)
  unsigned long long* bcur;      // [bands]: placement cursors
  unsigned long long* ew;        // [2m] entry weights, grouped by band
  uint16_t* eli;                 // [2m] entry index inside its band: (row - band * rows) * n + column
};

inline uint32_t band_rows(uint32_t n) { return n ? std::max(1u, std::min(n, kSpBandEntries / n)) : 1u; }

// Counts (kPlace false) or places (kPlace true) the entries of one tile of a
// node's edges: edge (u, v, w) gives entry (u, v) in band u / rows and (v, u) in
// band v / rows.  Placing reserves each band's range of the tile with one
// atomic on the band's cursor; the order inside a range is arbitrary, which
// the integer sums do not see.  Dynamic LDS: bands x (8 + 4) bytes.
template <bool kPlace>
__global__ void __launch_bounds__(kSpBucketBlock) k_sp_band_bucket(const SpDense* __restrict__ nodes) {
  extern __shared__ unsigned long long band_lds[];
  const SpDense d = nodes[blockIdx.y];
  const uint64_t e0 = (uint64_t)blockIdx.x * kSpBucketTile;
  if (e0 >= d.m) return;  // the same for the whole block
  unsigned long long* base = band_lds;                            // [bands] (place)
  uint32_t* cnt = (uint32_t*)(band_lds + (kPlace ? d.bands : 0));  // [bands]
  for (uint32_t b = threadIdx.x; b < d.bands; b += kSpBucketBlock) cnt[b] = 0;
  __syncthreads();
  const uint32_t R = d.rows;
  cac_edge x[kSpBucketEdges];
  uint32_t bu[kSpBucketEdges], bv[kSpBucketEdges];
#pragma unroll
  for (int j = 0; j < kSpBucketEdges; ++j) {
    const uint64_t e = e0 + (uint64_t)j * kSpBucketBlock + threadIdx.x;
    x[j] = e < d.m ? d.edges[e] : cac_edge{0u, 0u, 0ull};
  }
#pragma unroll
  for (int j = 0; j < kSpBucketEdges; ++j) {
    bu[j] = x[j].u / R;
    bv[j] = x[j].v / R;
    if (e0 + (uint64_t)j * kSpBucketBlock + threadIdx.x < d.m) {
      atomicAdd(&cnt[bu[j]], 1u);
      atomicAdd(&cnt[bv[j]], 1u);
    }
  }
  __syncthreads();
  if constexpr (!kPlace) {
    for (uint32_t b = threadIdx.x; b < d.bands; b += kSpBucketBlock)
      if (cnt[b]) atomicAdd(&d.boff[b], (unsigned long long)cnt[b]);
  } else {
    for (uint32_t b = threadIdx.x; b < d.bands; b += kSpBucketBlock) {
      base[b] = cnt[b] ? atomicAdd(&d.bcur[b], (unsigned long long)cnt[b]) : 0ull;
      cnt[b] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSpBucketEdges; ++j) {
      if (e0 + (uint64_t)j * kSpBucketBlock + threadIdx.x >= d.m) continue;
      const uint64_t pu = base[bu[j]] + atomicAdd(&cnt[bu[j]], 1u);
      d.ew[pu] = x[j].w;
      d.eli[pu] = (uint16_t)((x[j].u - bu[j] * R) * d.n + x[j].v);
      const uint64_t pv = base[bv[j]] + atomicAdd(&cnt[bv[j]], 1u);
      d.ew[pv] = x[j].w;
      d.eli[pv] = (uint16_t)((x[j].v - bv[j] * R) * d.n + x[j].u);
    }
  }
}

// Exclusive prefix of the band counts, one workgroup per node: boff[b] becomes
// the band's first entry, boff[bands] the total, and the cursors start there.
__global__ void __launch_bounds__(1024) k_sp_band_scan(const SpDense* __restrict__ nodes) {
  __shared__ unsigned long long wtot[16];
  const SpDense d = nodes[blockIdx.x];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long carry = 0;
  for (uint32_t b0 = 0; b0 < d.bands; b0 += 1024) {
    const uint32_t b = b0 + threadIdx.x;
    const unsigned long long c = b < d.bands ? d.boff[b] : 0ull;
    const unsigned long long inc = wave_incl_scan<uint64_t>(c);
    if (lane == 63) wtot[wave] = inc;
    __syncthreads();
    unsigned long long off = carry, tot = 0;
#pragma unroll
    for (int w = 0; w < 16; ++w) {
      if (w < wave) off += wtot[w];
      tot += wtot[w];
    }
    if (b < d.bands) {
      d.boff[b] = off + inc - c;
      d.bcur[b] = off + inc - c;
    }
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) d.boff[d.bands] = carry;
}

// Sums each band's entries in LDS, then writes the band's row sums and its
// rows (every entry, zeros included) and clears the LDS it read.  Workgroups
// stride over a node's bands.  Dynamic LDS: kSpBandEntries u64.
__global__ void __launch_bounds__(kSpBuildBlock) k_sp_band_build(const SpDense* __restrict__ nodes) {
  extern __shared__ unsigned long long band_acc[];
  const SpDense d = nodes[blockIdx.y];
  const uint32_t n = d.n, R = d.rows;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t i = threadIdx.x; i < R * n; i += kSpBuildBlock) band_acc[i] = 0;
  for (uint32_t b = blockIdx.x; b < d.bands; b += gridDim.x) {
    __syncthreads();  // LDS clear before the band's sums
    const uint64_t i1 = d.boff[b + 1];
    for (uint64_t i = d.boff[b] + threadIdx.x; i < i1; i += kSpBuildBlock) atomicAdd(&band_acc[d.eli[i]], d.ew[i]);
    __syncthreads();
    const uint32_t r0 = b * R, nr = min(R, n - r0);
    for (uint32_t r = wave; r < nr; r += kSpBuildBlock / 64) {
      uint64_t acc = 0;
      for (uint32_t c = lane; c < n; c += 64) acc += band_acc[r * n + c];
      acc = wave_sum<uint64_t>(acc);
      if (lane == 0) d.rsum[r0 + r] = (int64_t)acc;
    }
    __syncthreads();
    int64_t* out = d.mat + (uint64_t)r0 * n;
    for (uint32_t j = threadIdx.x; j < nr * n; j += kSpBuildBlock) {
      out[j] = (int64_t)band_acc[j];
      band_acc[j] = 0;
    }
  }
}

// Edge check (ids < n, u != v, 1 <= w < 2^40), u < v order, weighted degrees,
// and the total weight as sums of its high (w >> 20) and low 20-bit parts,
// which cannot wrap for fewer than 2^32 edges.  flags: bit 0 bad id or loop,
// bit 1 weight out of range.
__global__ void __launch_bounds__(256) k_sp_prep(cac_edge* e, uint64_t m, uint32_t n, unsigned long long* deg,
                                                 uint32_t* flags, unsigned long long* wsum) {
  uint32_t f = 0;
  unsigned long long whi = 0, wlo = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (uint64_t)gridDim.x * 256) {
    cac_edge x = e[i];
    if (x.u >= n || x.v >= n || x.u == x.v) {
      f |= 1;
      continue;
    }
    if (x.w == 0 || x.w >= (1ull << 40)) {
      f |= 2;
      continue;
    }
    if (x.u > x.v) {
      const uint32_t t = x.u;
      x.u = x.v;
      x.v = t;
      e[i] = x;
    }
    atomicAdd(&deg[x.u], (unsigned long long)x.w);
    atomicAdd(&deg[x.v], (unsigned long long)x.w);
    whi += x.w >> 20;
    wlo += x.w & 0xFFFFF;
  }
  if (f) atomicOr(flags, f);
  if (whi) atomicAdd(&wsum[0], whi);
  if (wlo) atomicAdd(&wsum[1], wlo);
}

__global__ void __launch_bounds__(256) k_sp_min(const unsigned long long* deg, uint32_t n, unsigned long long* out) {
  __shared__ unsigned long long w[4];
  unsigned long long b = ~0ull;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) b = min(b, deg[i]);
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) b = min(b, (unsigned long long)__shfl_xor(b, d, 64));
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = b;
  __syncthreads();
  if (threadIdx.x == 0) atomicMin(out, min(min(w[0], w[1]), min(w[2], w[3])));
}

__global__ void k_sp_fill(unsigned long long* p, uint64_t count, unsigned long long v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

inline bool hands_over(uint32_t n, uint64_t m) {
  return n <= kSparseDenseMax || (m >= (uint64_t)n * n / 32 && n <= kKsMaxS);
}

struct SpNode {
  const cac_edge* edges;
  uint64_t m;
  uint32_t n, trial, nid, slot;
  int64_t bound;  // the trial's bound (minimum weighted degree of its root)
  uint32_t seed;  // Philox key word 0 of the trial
};

// Device bytes of one child: what stays alive for the next depth (its edge
// list, descriptor, count) and the scratch it needs only while it contracts.
size_t child_keep_bytes(const SpNode& nd) { return carve_size(nd.m * sizeof(cac_edge)) + sizeof(SpChild) + 24; }
size_t child_tmp_bytes(const SpNode& nd) {
  return carve_size(nd.m * 8) + 2 * carve_size(nd.m * 4) + 2 * carve_size((size_t)nd.n * 4) +
         carve_size((nd.m + kSpTile - 1) / kSpTile * 4);
}
size_t dense_bytes(const SpNode& nd) {
  const uint32_t bands = (nd.n + band_rows(nd.n) - 1) / band_rows(nd.n);
  return carve_size((size_t)nd.n * nd.n * 8) + carve_size((size_t)nd.n * 8) + sizeof(SpDense) +
         ((size_t)bands + 1) * 8 + 256 + carve_size((size_t)bands * 8) + carve_size(nd.m * 16) + carve_size(nd.m * 4);
}

// The bucketing and band kernels' dynamic LDS may exceed the default limit:
// raise it once.
int band_attr(cac_ctx* ctx) {
  static std::once_flag once;
  static hipError_t err = hipSuccess;
  std::call_once(once, [] {
    err = hipFuncSetAttribute((const void*)k_sp_band_bucket<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kSpBucketLdsMax);
    if (err == hipSuccess)
      err = hipFuncSetAttribute((const void*)k_sp_band_bucket<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)kSpBucketLdsMax);
    if (err == hipSuccess)
      err = hipFuncSetAttribute((const void*)k_sp_band_build, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(kSpBandEntries * 8));
  });
  CAC_HIP(ctx, err);
  return CAC_OK;
}

struct SpRun {
  cac_ctx* ctx;
  int64_t* d_results;
  hipStream_t s;
  size_t depth_budget;  // bytes of children kept alive per depth
  size_t tmp_budget;    // bytes of contraction scratch per launch

  // Densifies nodes in batches that fit the free memory (matrices plus the
  // dense engine's level buffers) and runs the dense engine on them.
  int dense(const std::vector<SpNode>& nodes) {
    size_t i0 = 0;
    while (i0 < nodes.size()) {
      size_t avail = 0;
      int rc = device_memory_available(ctx, &avail);
      if (rc) return rc;
      avail += ctx->sp_dense_bytes;
      const size_t budget = (size_t)(0.6 * (double)avail);
      size_t i1 = i0, bytes = 256, mats = 256;
      uint32_t nmax = 0;
      while (i1 < nodes.size() && i1 - i0 < 4096) {
        const size_t b = dense_bytes(nodes[i1]) + ks_live_bytes_per_root(nodes[i1].n);
        if (i1 > i0 && bytes + b > budget) break;
        bytes += b;
        mats += dense_bytes(nodes[i1]);
        nmax = std::max(nmax, nodes[i1].n);
        ++i1;
      }
      const size_t count = i1 - i0;
      rc = grow_buffer(ctx, &ctx->sp_dense_buf, &ctx->sp_dense_bytes, mats);
      if (rc) return rc;
      rc = band_attr(ctx);
      if (rc) return rc;
      BufPin pin(ctx, kBufSpDense);
      Carver cv(ctx->sp_dense_buf);
      SpDense* d_nodes = cv.take<SpDense>(count);
      uint64_t nboff = 0;
      uint32_t bmax = 0;
      for (size_t i = 0; i < count; ++i) {
        const uint32_t n = nodes[i0 + i].n;
        if (n > kKsMaxS) return fail(ctx, CAC_EINVAL, "sparse densify: node of " + std::to_string(n) + " vertices");
        const uint32_t bands = (n + band_rows(n) - 1) / band_rows(n);
        nboff += bands + 1;
        bmax = std::max(bmax, bands);
      }
      unsigned long long* d_boff = cv.take<unsigned long long>(nboff);
      std::vector<SpDense> hd(count);
      std::vector<KsRoot> roots;
      uint64_t mmax = 0, boff_at = 0;
      for (size_t i = 0; i < count; ++i) {
        const SpNode& nd = nodes[i0 + i];
        const uint32_t R = band_rows(nd.n), bands = (nd.n + R - 1) / R;
        SpDense& d = hd[i];
        d.edges = nd.edges;
        d.m = nd.m;
        d.n = nd.n;
        d.rows = R;
        d.bands = bands;
        d.pad = 0;
        d.mat = cv.take<int64_t>((size_t)nd.n * nd.n);
        d.rsum = cv.take<int64_t>(nd.n);
        d.boff = d_boff + boff_at;
        boff_at += bands + 1;
        d.bcur = cv.take<unsigned long long>(bands);
        d.ew = cv.take<unsigned long long>(2 * nd.m);
        d.eli = cv.take<uint16_t>(2 * nd.m);
        KsRoot r{d.mat, d.rsum, nd.n, nd.trial, nd.slot, nd.bound, nd.seed};
        r.nid = nd.nid;
        roots.push_back(r);
        mmax = std::max(mmax, nd.m);
      }
      CAC_HIP(ctx, hipMemcpyAsync(d_nodes, hd.data(), count * sizeof(SpDense), hipMemcpyHostToDevice, s));
      const unsigned tiles = (unsigned)std::max<uint64_t>(1, (mmax + kSpBucketTile - 1) / kSpBucketTile);
      const unsigned gb = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(bmax, (2048 + count - 1) / count));
      prof_begin(ctx, "sparse_densify", s);
      hipLaunchKernelGGL(k_sp_fill, dim3(grid_for(nboff, 256, 1024)), dim3(256), 0, s, d_boff, nboff, 0ull);
      hipLaunchKernelGGL(k_sp_band_bucket<false>, dim3(tiles, (unsigned)count), dim3(kSpBucketBlock),
                         (size_t)bmax * 4, s, (const SpDense*)d_nodes);
      hipLaunchKernelGGL(k_sp_band_scan, dim3((unsigned)count), dim3(1024), 0, s, (const SpDense*)d_nodes);
      hipLaunchKernelGGL(k_sp_band_bucket<true>, dim3(tiles, (unsigned)count), dim3(kSpBucketBlock),
                         (size_t)bmax * 12, s, (const SpDense*)d_nodes);
      hipLaunchKernelGGL(k_sp_band_build, dim3(gb, (unsigned)count), dim3(kSpBuildBlock), (size_t)kSpBandEntries * 8,
                         s, (const SpDense*)d_nodes);
      prof_end(ctx, s);
      CAC_LAUNCH_CHECK(ctx, "sparse densify");
      rc = ks_run(ctx, roots, d_results, s);
      if (rc) return rc;
      i0 = i1;
    }
    return CAC_OK;
  }

  // One depth of the recursion: hands over what is small or dense enough, and
  // contracts the rest in chunks whose children fit this depth's buffer, each
  // chunk's subtree finished before the next chunk reuses the buffer.
  int level(const std::vector<SpNode>& nodes, size_t depth) {
    std::vector<SpNode> dn, sp;
    for (const SpNode& nd : nodes) (hands_over(nd.n, nd.m) ? dn : sp).push_back(nd);
    if (!dn.empty()) {
      int rc = dense(dn);
      if (rc) return rc;
    }
    if (sp.empty()) return CAC_OK;
    if (ctx->sp_bufs.size() <= depth) {
      ctx->sp_bufs.resize(depth + 1, nullptr);
      ctx->sp_bytes.resize(depth + 1, 0);
    }
    size_t i0 = 0;
    while (i0 < sp.size()) {
      size_t i1 = i0, bytes = 1024, tmp = 1024;
      while (i1 < sp.size() && i1 - i0 < 8192) {
        const size_t b = 2 * child_keep_bytes(sp[i1]), t = 2 * child_tmp_bytes(sp[i1]);
        if (i1 > i0 && (bytes + b > depth_budget || tmp + t > tmp_budget)) break;
        bytes += b;
        tmp += t;
        ++i1;
      }
      const size_t count = i1 - i0, nchild = 2 * count;
      int rc = grow_buffer(ctx, &ctx->sp_bufs[depth], &ctx->sp_bytes[depth], bytes);
      if (rc) return rc;
      BufPin pin(ctx, kBufSpDepth + (int)depth);  // this chunk's children, until their subtree is done
      rc = grow_buffer(ctx, &ctx->sp_tmp, &ctx->sp_tmp_bytes, tmp);
      if (rc) return rc;
      Carver cv(ctx->sp_bufs[depth]);
      Carver ct(ctx->sp_tmp);
      SpChild* d_children = cv.take<SpChild>(nchild);
      unsigned long long* d_mout = cv.take<unsigned long long>(nchild);
      unsigned long long* d_kmm = cv.take<unsigned long long>(2 * nchild);
      std::vector<SpChild> ch(nchild);
      uint64_t mmax = 0;
      for (size_t i = 0; i < count; ++i) {
        const SpNode& nd = sp[i0 + i];
        for (uint32_t k = 0; k < 2; ++k) {
          SpChild& x = ch[2 * i + k];
          x.edges = nd.edges;
          x.m = nd.m;
          x.n = nd.n;
          x.tgt = ks_target_host(nd.n);
          x.trial = nd.trial;
          x.nid = 2 * nd.nid + k;
          x.slot = nd.slot;
          x.seed = nd.seed;
          x.out = cv.take<cac_edge>(nd.m);
          x.keys = ct.take<uint64_t>(nd.m);
          x.list[0] = ct.take<uint32_t>(nd.m);
          x.list[1] = ct.take<uint32_t>(nd.m);
          x.parc = ct.take<uint32_t>(nd.n);
          x.part = ct.take<uint32_t>(nd.n);
          x.tiles = ct.take<uint32_t>((nd.m + kSpTile - 1) / kSpTile);
          x.m_out = d_mout + 2 * i + k;
          x.kmm = d_kmm + 2 * (2 * i + k);
        }
        mmax = std::max<uint64_t>(mmax, nd.m);
      }
      CAC_HIP(ctx, hipMemcpyAsync(d_children, ch.data(), nchild * sizeof(SpChild), hipMemcpyHostToDevice, s));
      prof_begin(ctx, "sparse_contract", s);
      launch_sparse_contract(d_children, (uint32_t)nchild, mmax, d_results, s);
      prof_end(ctx, s);
      CAC_LAUNCH_CHECK(ctx, "sparse contract");
      std::vector<unsigned long long> mout(nchild);
      CAC_HIP(ctx, hipMemcpyAsync(mout.data(), d_mout, nchild * 8, hipMemcpyDeviceToHost, s));
      CAC_HIP(ctx, hipStreamSynchronize(s));
      std::vector<SpNode> next;
      for (size_t i = 0; i < nchild; ++i) {
        if (mout[i] == kSpDisconnected) continue;  // its cut of 0 is already recorded
        next.push_back(
            {ch[i].out, mout[i], ch[i].tgt, ch[i].trial, ch[i].nid, ch[i].slot, sp[i0 + i / 2].bound, ch[i].seed});
      }
      rc = level(next, depth + 1);
      if (rc) return rc;
      i0 = i1;
    }
    return CAC_OK;
  }
};

// Depth of the sparse recursion at most: until the node size alone hands over.
size_t sparse_depth(uint32_t n) {
  size_t d = 0;
  for (uint32_t s = n; s > kSparseDenseMax; s = ks_target_host(s)) ++d;
  return d;
}

}  // namespace

int ks_sparse_roots(cac_ctx* ctx, const std::vector<SparseRoot>& roots, int64_t* d_results, hipStream_t s) {
  if (roots.empty()) return CAC_OK;
  size_t avail = 0;
  int rc = device_memory_available(ctx, &avail);
  if (rc) return rc;
  for (size_t i = 0; i < ctx->sp_bufs.size(); ++i) avail += ctx->sp_bytes[i];
  avail += ctx->sp_dense_bytes + ctx->sp_tmp_bytes;
  uint32_t nmax = 0;
  for (const SparseRoot& r : roots) nmax = std::max(nmax, r.n);
  const size_t depth = sparse_depth(nmax);
  SpRun run{ctx, d_results, s, (size_t)(0.4 * (double)avail) / (depth + 1), (size_t)(0.15 * (double)avail)};
  std::vector<SpNode> level;
  for (const SparseRoot& r : roots) level.push_back({r.edges, r.m, r.n, r.trial, 1u, r.slot, r.bound, r.seed});
  rc = run.level(level, 0);
  if (rc) return rc;
  CAC_HIP(ctx, hipStreamSynchronize(s));
  return CAC_OK;
}

int ks_sparse_run(cac_ctx* ctx, const cac_edge* d_edges, uint64_t m, uint32_t n, uint32_t seed,
                  const std::vector<uint32_t>& trial_ids, int64_t bound, int64_t* d_results, hipStream_t s) {
  std::vector<SparseRoot> roots;
  for (size_t i = 0; i < trial_ids.size(); ++i)
    roots.push_back({d_edges, m, n, trial_ids[i], (uint32_t)i, bound, seed});
  return ks_sparse_roots(ctx, roots, d_results, s);
}

}  // namespace cac

extern "C" {

uint64_t cac_race_key(uint64_t x, uint64_t w) { return cac::race_key(x, w); }

int cac_ks_sparse(cac_ctx* ctx, const cac_edge* edges, uint64_t m, uint32_t n, uint32_t seed, uint32_t first_trial,
                  uint32_t ntrials, int64_t* raw_cuts, int64_t* min_cut) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if (!edges || m == 0 || n < 2 || ntrials == 0 || !min_cut)
    return cac::fail(ctx, CAC_EINVAL, "need n >= 2, at least one edge and one trial");
  if (m >= 0xFFFFFFFFull) return cac::fail(ctx, CAC_EINVAL, "sparse Karger-Stein supports fewer than 2^32 - 1 edges");
  hipStream_t s = ctx->stream;
  cac_edge* d_e = nullptr;
  unsigned long long* d_aux = nullptr;  // degrees [n], minimum, flags, weight sum (high, low parts)
  int64_t* d_res = nullptr;
  hipError_t e = hipMalloc((void**)&d_e, m * sizeof(cac_edge));
  if (e == hipSuccess) e = hipMalloc((void**)&d_aux, 8 * ((size_t)n + 4));
  if (e == hipSuccess) e = hipMalloc((void**)&d_res, 8 * (size_t)ntrials);
  if (e != hipSuccess) rc = cac::hip_fail(ctx, e, "sparse Karger-Stein buffers");
  uint32_t flags = 0;
  unsigned long long bound = 0, wsum[2] = {0, 0};
  if (rc == CAC_OK) {
    unsigned long long* d_min = d_aux + n;
    uint32_t* d_flags = (uint32_t*)(d_aux + n + 1);
    unsigned long long* d_wsum = d_aux + n + 2;
    e = hipMemcpyAsync(d_e, edges, m * sizeof(cac_edge), hipMemcpyHostToDevice, s);
    hipLaunchKernelGGL(cac::k_sp_fill, dim3(cac::grid_for((uint64_t)n + 4, 256, 1024)), dim3(256), 0, s, d_aux,
                       (uint64_t)n + 4, 0ull);
    hipLaunchKernelGGL(cac::k_sp_fill, dim3(1), dim3(64), 0, s, d_min, (uint64_t)1, ~0ull);
    hipLaunchKernelGGL(cac::k_sp_prep, dim3(cac::grid_for(m, 256, 4096)), dim3(256), 0, s, d_e, m, n, d_aux,
                       d_flags, d_wsum);
    hipLaunchKernelGGL(cac::k_sp_min, dim3(cac::grid_for(n, 256, 1024)), dim3(256), 0, s,
                       (const unsigned long long*)d_aux, n, d_min);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&bound, d_min, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&flags, d_flags, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(wsum, d_wsum, 16, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = cac::hip_fail(ctx, e, "sparse Karger-Stein edge check");
  }
  if (rc == CAC_OK && (flags & 1)) rc = cac::fail(ctx, CAC_EINVAL, "edge endpoint >= n, or a loop");
  if (rc == CAC_OK && (flags & 2)) rc = cac::fail(ctx, CAC_ERANGE, "edge weight must be in [1, 2^40)");
  if (rc == CAC_OK && ((unsigned __int128)wsum[0] << 20) + wsum[1] >= ((unsigned __int128)1 << 61))
    rc = cac::fail(ctx, CAC_ERANGE, "total edge weight must stay below 2^61");
  if (rc == CAC_OK) {
    hipLaunchKernelGGL(cac::k_sp_fill, dim3(cac::grid_for(ntrials, 256, 64)), dim3(256), 0, s,
                       (unsigned long long*)d_res, (uint64_t)ntrials, bound);
    std::vector<uint32_t> ids(ntrials);
    for (uint32_t i = 0; i < ntrials; ++i) ids[i] = first_trial + i;
    rc = cac::ks_sparse_run(ctx, d_e, m, n, seed, ids, (int64_t)bound, d_res, s);
  }
  if (rc == CAC_OK) {
    std::vector<int64_t> res(ntrials);
    e = hipMemcpyAsync(res.data(), d_res, 8 * (size_t)ntrials, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = cac::hip_fail(ctx, e, "sparse Karger-Stein results");
    int64_t mn = INT64_MAX;
    for (uint32_t t = 0; t < ntrials && rc == CAC_OK; ++t) {
      if (raw_cuts) raw_cuts[t] = res[t];
      mn = std::min(mn, res[t]);
    }
    if (rc == CAC_OK) *min_cut = mn;
  }
  (void)hipStreamSynchronize(s);
  if (d_res) (void)hipFree(d_res);
  if (d_aux) (void)hipFree(d_aux);
  if (d_e) (void)hipFree(d_e);
  return rc;
}

int cac_sparse_contract(cac_ctx* ctx, const cac_edge* edges, uint64_t m, uint32_t n, uint32_t tgt, uint32_t seed,
                        uint32_t trial, uint32_t nid, cac_edge* out, uint64_t out_cap, uint64_t* m_out) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if (!edges || !m_out || (m > 0 && !out) || n < 1 || tgt < 1)
    return cac::fail(ctx, CAC_EINVAL, "need edges, n >= 1, tgt >= 1 and an output");
  if (m >= 0xFFFFFFFFull) return cac::fail(ctx, CAC_EINVAL, "sparse contraction supports fewer than 2^32 - 1 edges");
  for (uint64_t i = 0; i < m; ++i) {
    const cac_edge& x = edges[i];
    if (x.u >= n || x.v >= n || x.u >= x.v) return cac::fail(ctx, CAC_EINVAL, "edges must have u < v < n");
    if (x.w == 0 || x.w >= (1ull << 40)) return cac::fail(ctx, CAC_ERANGE, "edge weight must be in [1, 2^40)");
  }
  hipStream_t s = ctx->stream;
  const size_t mm = m ? m : 1;
  const size_t bytes = cac::carve_size(sizeof(cac::SpChild)) + cac::carve_size(mm * sizeof(cac_edge)) * 2 +
                       cac::carve_size(mm * 8) + 2 * cac::carve_size(mm * 4) + 2 * cac::carve_size((size_t)n * 4) +
                       cac::carve_size(16) + cac::carve_size(16) + cac::carve_size((mm + cac::kSpTile - 1) / cac::kSpTile * 4) +
                       256;
  void* base = nullptr;
  rc = cac::scratch_reserve(ctx, bytes, &base);
  if (rc) return rc;
  cac::Carver cv(base);
  cac::SpChild c{};
  cac::SpChild* d_child = cv.take<cac::SpChild>(1);
  cac_edge* d_in = cv.take<cac_edge>(mm);
  c.edges = d_in;
  c.m = m;
  c.n = n;
  c.tgt = tgt;
  c.trial = trial;
  c.nid = nid;
  c.seed = seed;
  c.out = cv.take<cac_edge>(mm);
  c.keys = cv.take<uint64_t>(mm);
  c.list[0] = cv.take<uint32_t>(mm);
  c.list[1] = cv.take<uint32_t>(mm);
  c.parc = cv.take<uint32_t>(n);
  c.part = cv.take<uint32_t>(n);
  unsigned long long* d_mout = cv.take<unsigned long long>(1);
  int64_t* d_res = cv.take<int64_t>(1);
  c.m_out = d_mout;
  c.kmm = cv.take<unsigned long long>(2);
  c.tiles = cv.take<uint32_t>((mm + cac::kSpTile - 1) / cac::kSpTile);
  if (m) CAC_HIP(ctx, hipMemcpyAsync(d_in, edges, m * sizeof(cac_edge), hipMemcpyHostToDevice, s));
  CAC_HIP(ctx, hipMemcpyAsync(d_child, &c, sizeof(c), hipMemcpyHostToDevice, s));
  cac::launch_sparse_contract(d_child, 1, m, d_res, s);
  CAC_LAUNCH_CHECK(ctx, "sparse contract");
  unsigned long long mo = 0;
  CAC_HIP(ctx, hipMemcpyAsync(&mo, d_mout, 8, hipMemcpyDeviceToHost, s));
  CAC_HIP(ctx, hipStreamSynchronize(s));
  if (mo == cac::kSpDisconnected) return cac::fail(ctx, CAC_ENOTCONN, "more than tgt components");
  *m_out = mo;
  if (mo > out_cap) return cac::fail(ctx, CAC_ERANGE, "output capacity too small");
  if (mo) CAC_HIP(ctx, hipMemcpy(out, c.out, mo * sizeof(cac_edge), hipMemcpyDeviceToHost));
  return CAC_OK;
}

int cac_karger_stein(cac_ctx* ctx, const cac_edge* edges, uint64_t m, uint32_t n, double p, uint32_t seed,
                     int64_t* cut, uint32_t* trials) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if (!edges || m == 0 || n < 2 || !cut) return cac::fail(ctx, CAC_EINVAL, "need n >= 2 and at least one edge");
  if (!(p >= 0.0 && p < 1.0)) return cac::fail(ctx, CAC_EINVAL, "success probability must be in [0, 1)");
  const uint32_t T = cac::ks_trials_host(n, p);
  if (trials) *trials = T;
  // the reference's dense switch (co_mincut.h:401); a dense graph larger than
  // the dense engine's 8192 vertices starts in the sparse recursion, which
  // hands it over once a node is small enough
  if (m >= (uint64_t)n * n / 32 && n <= cac::kKsMaxS) {
    std::vector<int64_t> A((size_t)n * n, 0);
    unsigned __int128 total = 0;
    for (uint64_t i = 0; i < m; ++i) {
      const cac_edge& x = edges[i];
      if (x.u >= n || x.v >= n || x.u == x.v) return cac::fail(ctx, CAC_EINVAL, "edge endpoint >= n, or a loop");
      if (x.w == 0 || x.w >= (1ull << 40)) return cac::fail(ctx, CAC_ERANGE, "edge weight must be in [1, 2^40)");
      total += x.w;
      if (total >= ((unsigned __int128)1 << 61)) return cac::fail(ctx, CAC_ERANGE, "total edge weight must stay below 2^61");
      A[(size_t)x.u * n + x.v] += (int64_t)x.w;
      A[(size_t)x.v * n + x.u] += (int64_t)x.w;
    }
    return cac_ks_dense(ctx, A.data(), n, seed, 0, T, nullptr, cut);
  }
  return cac_ks_sparse(ctx, edges, m, n, seed, 0, T, nullptr, cut);
}

}  // extern "C"
