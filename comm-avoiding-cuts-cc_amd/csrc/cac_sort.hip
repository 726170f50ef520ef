// cac_sort.hip -- LDS-staged LSD radix sort of (u64 key, u64 value) pairs and
// the segmented integer sum of equal keys, for the merge of a contracted edge
// list (ranked endpoints -> keys, parallel edges -> summed weights).
//
// Replaces the merge step of the reference's contraction: the sort of the
// relabelled edges and the sum of parallel weights in
// AdjacencyListGraph::finalizeContractionPhase (src/AdjacencyListGraph.cpp:87-112)
// and WeightedIteratedSparseSampling::reduce with its SamplingSorter
// (src/WeightedIteratedSparseSampling.cpp:26-234, src/sorting/SamplingSorter.hpp:44-180).
//
// One pass per 8-bit digit, three launches:
//   1. k_radix_hist: per 2048-pair tile, the digit histogram in LDS.  Lanes of
//      a wavefront holding equal digits are found by eight ballots and one
//      lane adds their count, so a skewed digit (a constant trial field)
//      costs one LDS atomic per wavefront instead of one per pair.
//   2. a device-wide exclusive scan of the tile histograms in digit-major
//      order: the global start of every (digit, tile) run.
//   3. k_radix_scatter: the tile again, every pair ranked within its
//      wavefront by the same ballot match (stable: wavefront slices and rounds
//      follow the input order), the wavefronts' counts scanned per digit, the
//      tile's pairs placed in LDS in digit order, then written out from LDS so
//      that consecutive lanes write consecutive addresses of one digit's run.
// Integer work, HBM-bound: each pass reads and writes 16 bytes per pair.
#include <cstdint>

#include "cac_common.h"
#include "cac_internal.h"
#include "cac_scan.h"
#include "cac_sort.h"
#include "cac_wave.h"

namespace cac {
namespace {

constexpr int kSortBlock = 256;
constexpr int kSortWaves = kSortBlock / 64;
constexpr int kSortItems = 8;                          // pairs per thread
constexpr int kSortTile = kSortBlock * kSortItems;     // 2048 pairs per tile
constexpr int kWaveSlice = 64 * kSortItems;            // pairs per wavefront slice
constexpr int kBins = 256;
static_assert(kBins == kSortBlock, "one digit per thread in the per-digit phases");

__device__ inline uint32_t digit_of(uint64_t key, int shift) { return (uint32_t)(key >> shift) & (kBins - 1); }

// Lanes (among `live`) whose digit equals this lane's.
__device__ inline uint64_t match_digit(uint32_t d, uint64_t live) {
  uint64_t m = live;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const uint64_t bal = __ballot((d >> b) & 1u);
    m &= ((d >> b) & 1u) ? bal : ~bal;
  }
  return m;
}

__device__ inline uint64_t lanes_below() {
  const int lane = threadIdx.x & 63;
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// counts[d * ntiles + tile] = pairs of the tile with digit d.
__global__ void __launch_bounds__(kSortBlock) k_radix_hist(const uint64_t* __restrict__ keys, uint64_t n, int shift,
                                                           uint32_t* __restrict__ counts, uint32_t ntiles) {
  __shared__ uint32_t h[kBins];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t base = (uint64_t)blockIdx.x * kSortTile + (uint64_t)wave * kWaveSlice;
#pragma unroll 4
  for (int r = 0; r < kSortItems; ++r) {
    const uint64_t i = base + (uint64_t)r * 64 + lane;
    const bool ok = i < n;
    const uint32_t d = ok ? digit_of(keys[i], shift) : 0u;
    const uint64_t m = match_digit(d, __ballot(ok));
    if (ok && (m & lanes_below()) == 0) atomicAdd(&h[d], (uint32_t)__popcll(m));
  }
  __syncthreads();
  counts[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// One tile: ranks, LDS staging in digit order, coalesced writes.
__global__ void __launch_bounds__(kSortBlock) k_radix_scatter(const uint64_t* __restrict__ keys,
                                                              const uint64_t* __restrict__ vals, uint64_t n, int shift,
                                                              const uint32_t* __restrict__ offsets, uint32_t ntiles,
                                                              uint64_t* __restrict__ kout, uint64_t* __restrict__ vout) {
  __shared__ uint32_t wcnt[kSortWaves][kBins];  // per wavefront digit counts, then their exclusive prefix
  __shared__ uint32_t tstart[kBins];            // digit start within the tile
  __shared__ uint32_t gbase[kBins];             // digit start in the output
  __shared__ uint32_t wsum[kSortWaves];
  __shared__ uint64_t skey[kSortTile];
  __shared__ uint64_t sval[kSortTile];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int w = 0; w < kSortWaves; ++w) wcnt[w][threadIdx.x] = 0;
  gbase[threadIdx.x] = offsets[(uint64_t)threadIdx.x * ntiles + blockIdx.x];
  __syncthreads();
  const uint64_t tile0 = (uint64_t)blockIdx.x * kSortTile;
  const uint64_t base = tile0 + (uint64_t)wave * kWaveSlice;
  uint64_t k[kSortItems], v[kSortItems];
  uint32_t off[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const uint64_t i = base + (uint64_t)r * 64 + lane;
    k[r] = i < n ? keys[i] : 0ull;
    v[r] = i < n ? vals[i] : 0ull;
  }
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const bool ok = base + (uint64_t)r * 64 + lane < n;
    const uint32_t d = digit_of(k[r], shift);
    const uint64_t m = match_digit(d, __ballot(ok));
    const uint32_t before = (uint32_t)__popcll(m & lanes_below());
    off[r] = wcnt[wave][d] + before;  // every lane reads before the group's leader adds
    if (ok && before == 0) wcnt[wave][d] += (uint32_t)__popcll(m);
  }
  __syncthreads();
  // per digit (one per thread): exclusive prefix over the wavefronts, then
  // over the digits for the tile
  uint32_t tot = 0;
  for (int w = 0; w < kSortWaves; ++w) {
    const uint32_t c = wcnt[w][threadIdx.x];
    wcnt[w][threadIdx.x] = tot;
    tot += c;
  }
  {
    const uint32_t inc = wave_incl_scan<uint32_t>(tot);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < wave; ++w) add += wsum[w];
    tstart[threadIdx.x] = add + inc - tot;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    if (base + (uint64_t)r * 64 + lane >= n) continue;
    const uint32_t d = digit_of(k[r], shift);
    const uint32_t p = tstart[d] + wcnt[wave][d] + off[r];
    skey[p] = k[r];
    sval[p] = v[r];
  }
  __syncthreads();
  const uint32_t count = (uint32_t)(n - tile0 < (uint64_t)kSortTile ? n - tile0 : (uint64_t)kSortTile);
  for (uint32_t j = threadIdx.x; j < count; j += kSortBlock) {
    const uint64_t key = skey[j];
    const uint32_t d = digit_of(key, shift);
    const uint64_t g = (uint64_t)gbase[d] + (j - tstart[d]);
    kout[g] = key;
    vout[g] = sval[j];
  }
}

struct CountLoad {
  const uint32_t* c;
  __device__ uint32_t operator()(uint64_t i) const { return c[i]; }
};
struct CountStore {
  uint32_t* c;
  __device__ void operator()(uint64_t i, uint32_t ex, uint32_t) const { c[i] = ex; }
};

// ---- segmented sum --------------------------------------------------------
struct HeadFlag {
  const uint64_t* keys;
  uint64_t drop;
  __device__ uint32_t operator()(uint64_t i) const {
    const uint64_t k = keys[i];
    return ((k & drop) != drop && (i == 0 || keys[i - 1] != k)) ? 1u : 0u;
  }
};
struct HeadStore {
  uint32_t* idx;
  __device__ void operator()(uint64_t i, uint32_t ex, uint32_t) const { idx[i] = ex; }
};
struct ValLoad {
  const uint64_t* v;
  __device__ uint64_t operator()(uint64_t i) const { return v[i]; }
};
struct ValStore {
  uint64_t* e;
  __device__ void operator()(uint64_t i, uint64_t ex, uint64_t) const { e[i] = ex; }
};

// Run heads: the key and the value prefix before the run.
__global__ void __launch_bounds__(256) k_seg_heads(const uint64_t* __restrict__ keys, const uint64_t* __restrict__ pre,
                                                   uint64_t n, uint64_t drop, const uint32_t* __restrict__ idx,
                                                   uint64_t* __restrict__ okeys, uint64_t* __restrict__ ovals,
                                                   uint32_t stride) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const uint64_t k = keys[i];
    if ((k & drop) == drop || (i > 0 && keys[i - 1] == k)) continue;
    okeys[(uint64_t)idx[i] * stride] = k;
    ovals[(uint64_t)idx[i] * stride] = pre[i];
  }
}

// Run tails: the sum = prefix through the tail - prefix before the head.
__global__ void __launch_bounds__(256) k_seg_tails(const uint64_t* __restrict__ keys, const uint64_t* __restrict__ vals,
                                                   const uint64_t* __restrict__ pre, uint64_t n, uint64_t drop,
                                                   const uint32_t* __restrict__ idx, uint64_t* __restrict__ ovals,
                                                   uint32_t stride) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const uint64_t k = keys[i];
    if ((k & drop) == drop || (i + 1 < n && keys[i + 1] == k)) continue;
    const uint64_t j = (uint64_t)(idx[i + 1] - 1) * stride;  // this run's output slot
    ovals[j] = pre[i] + vals[i] - ovals[j];
  }
}

}  // namespace

size_t radix_sort_temp_bytes(uint64_t n) {
  const uint64_t ntiles = (n + kSortTile - 1) / kSortTile;
  const uint64_t cells = ntiles * kBins;
  return carve_size(cells * 4) + carve_size((scan_tiles(cells) + 64) * 4) + 256;
}

int radix_sort_pairs(cac_ctx* ctx, uint64_t* keys, uint64_t* vals, uint64_t* kalt, uint64_t* valt, uint64_t n,
                     int bits, void* temp, bool* in_alt, hipStream_t s) {
  *in_alt = false;
  if (n == 0 || bits <= 0) return CAC_OK;
  if (n >= 0xFFFFFFFFull) return fail(ctx, CAC_EINVAL, "radix sort supports fewer than 2^32 - 1 pairs");
  const uint32_t ntiles = (uint32_t)((n + kSortTile - 1) / kSortTile);
  const uint64_t cells = (uint64_t)ntiles * kBins;
  Carver cv(temp);
  uint32_t* counts = cv.take<uint32_t>(cells);
  uint32_t* partial = cv.take<uint32_t>(scan_tiles(cells) + 64);
  uint64_t *ki = keys, *vi = vals, *ko = kalt, *vo = valt;
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(kSortBlock), 0, s, (const uint64_t*)ki, n, shift, counts,
                       ntiles);
    hipError_t e = device_exclusive_scan<uint32_t>(CountLoad{counts}, CountStore{counts}, cells, partial,
                                                   (uint32_t*)nullptr, s);
    if (e != hipSuccess) return hip_fail(ctx, e, "radix sort scan");
    hipLaunchKernelGGL(k_radix_scatter, dim3(ntiles), dim3(kSortBlock), 0, s, (const uint64_t*)ki,
                       (const uint64_t*)vi, n, shift, (const uint32_t*)counts, ntiles, ko, vo);
    std::swap(ki, ko);
    std::swap(vi, vo);
    *in_alt = !*in_alt;
  }
  CAC_LAUNCH_CHECK(ctx, "radix sort");
  return CAC_OK;
}

size_t segmented_sum_temp_bytes(uint64_t n) {
  return carve_size(std::max<uint64_t>(n, 1) * 8) + carve_size((scan_tiles(n) + 64) * 8) + 256;
}

int segmented_sum(cac_ctx* ctx, const uint64_t* keys, const uint64_t* vals, uint64_t n, uint64_t drop_mask,
                  uint64_t* out_keys, uint64_t* out_vals, uint32_t stride, uint32_t* head_idx, void* temp,
                  hipStream_t s) {
  if (n >= 0xFFFFFFFFull) return fail(ctx, CAC_EINVAL, "segmented sum supports fewer than 2^32 - 1 pairs");
  Carver cv(temp);
  uint64_t* pre = cv.take<uint64_t>(std::max<uint64_t>(n, 1));
  uint64_t* partial = cv.take<uint64_t>(scan_tiles(n) + 64);
  hipError_t e = device_exclusive_scan<uint32_t>(HeadFlag{keys, drop_mask}, HeadStore{head_idx}, n,
                                                 reinterpret_cast<uint32_t*>(partial), head_idx + n, s);
  if (e == hipSuccess) e = device_exclusive_scan<uint64_t>(ValLoad{vals}, ValStore{pre}, n, partial,
                                                           (uint64_t*)nullptr, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "segmented sum scans");
  if (n) {
    const unsigned g = grid_for(n, 256, 8192);
    hipLaunchKernelGGL(k_seg_heads, dim3(g), dim3(256), 0, s, keys, (const uint64_t*)pre, n, drop_mask,
                       (const uint32_t*)head_idx, out_keys, out_vals, stride);
    hipLaunchKernelGGL(k_seg_tails, dim3(g), dim3(256), 0, s, keys, vals, (const uint64_t*)pre, n, drop_mask,
                       (const uint32_t*)head_idx, out_vals, stride);
  }
  CAC_LAUNCH_CHECK(ctx, "segmented sum");
  return CAC_OK;
}

}  // namespace cac
