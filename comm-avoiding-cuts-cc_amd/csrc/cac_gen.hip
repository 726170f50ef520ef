// cac_gen.hip -- synthetic inputs generated on the device (R-MAT edge lists).
//
// The reference builds its R-MAT inputs offline with PaRMAT
// (input_generators/rmat_driver.sh:13-16) and reads them back as text.  Here the
// graph is produced directly in HBM from a counter-based stream so a benchmark
// input of 2^28 edges costs milliseconds instead of a 5 GB text parse.  Sorting
// uses rocPRIM's radix sort: the generator is set-up work, not the measured path.
#include <cmath>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "cac_common.h"
#include "cac_internal.h"
#include "cac_scan.h"

namespace cac {
namespace {

__global__ void __launch_bounds__(256) k_rmat_keys(uint64_t draws, uint32_t scale, uint32_t seed, uint32_t ta,
                                                   uint32_t tab, uint32_t tabc, uint64_t* keys, uint32_t* vals) {
  const uint64_t loop_key = (scale == 32) ? ~0ull : ((1ull << (2 * scale)) - 1);
  for (uint64_t d = (uint64_t)blockIdx.x * 256 + threadIdx.x; d < draws; d += (uint64_t)gridDim.x * 256) {
    uint32_t u = 0, v = 0;
    u32x4 w{0, 0, 0, 0};
    for (uint32_t lvl = 0; lvl < scale; ++lvl) {
      if ((lvl & 3u) == 0) w = philox4x32((uint32_t)d, lvl >> 2, (uint32_t)(d >> 32), 0u, seed, kDomRmat);
      uint32_t r = (lvl & 3u) == 0 ? w.x : (lvl & 3u) == 1 ? w.y : (lvl & 3u) == 2 ? w.z : w.w;
      uint32_t bu = r >= tab ? 1u : 0u;
      uint32_t bv = (r >= ta && r < tab) || r >= tabc ? 1u : 0u;
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    uint64_t key;
    if (u == v) key = loop_key;
    else {
      uint32_t a = u < v ? u : v, b = u < v ? v : u;
      key = ((uint64_t)a << scale) | b;
    }
    keys[d] = key;
    vals[d] = (uint32_t)d;
  }
}

// first[i]: key i is the first (smallest-draw) copy of its pair; mark[draw] = first.
__global__ void __launch_bounds__(256) k_rmat_first(const uint64_t* keys, const uint32_t* vals, uint64_t draws,
                                                    uint64_t loop_key, uint8_t* mark, uint8_t* first) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < draws; i += (uint64_t)gridDim.x * 256) {
    uint64_t k = keys[i];
    uint8_t f = (k != loop_key && (i == 0 || keys[i - 1] != k)) ? 1 : 0;
    first[i] = f;
    mark[vals[i]] = f;
  }
}

struct ByteLoad {
  const uint8_t* p;
  __device__ uint32_t operator()(uint64_t i) const { return p[i]; }
};
// Finds the draw whose first occurrence completes the m-th distinct pair.
struct CutoffStore {
  uint64_t m;
  uint64_t* cutoff;
  __device__ void operator()(uint64_t i, uint32_t ex, uint32_t v) const {
    if (v && (uint64_t)ex + 1 == m) *cutoff = i;
  }
};
struct KeepLoad {
  const uint8_t* first;
  const uint32_t* vals;
  const uint64_t* cutoff;
  __device__ uint32_t operator()(uint64_t i) const { return (first[i] && vals[i] <= *cutoff) ? 1u : 0u; }
};
struct EmitStore {
  const uint64_t* keys;
  uint32_t scale;
  cac_uedge* out;
  __device__ void operator()(uint64_t i, uint32_t ex, uint32_t v) const {
    if (!v) return;
    uint64_t k = keys[i];
    out[ex] = cac_uedge{(uint32_t)(k >> scale), (uint32_t)(k & ((1ull << scale) - 1))};
  }
};

__global__ void k_init_cutoff(uint64_t* c) { *c = ~0ull; }

}  // namespace
}  // namespace cac

extern "C" int cac_gen_rmat_dev(cac_ctx* ctx, uint32_t scale, uint64_t m, uint32_t seed, uint32_t ta, uint32_t tab,
                                uint32_t tabc, cac_uedge* d_edges, uint64_t* draws_used) {
  using namespace cac;
  int rc = bind(ctx);
  if (rc) return rc;
  if (scale < 1 || scale > 31 || !(ta <= tab && tab <= tabc)) return fail(ctx, CAC_EINVAL, "bad R-MAT parameters");
  if (m && !d_edges) return fail(ctx, CAC_EINVAL, "null edges");
  const uint64_t nv = 1ull << scale;
  if (m > nv * (nv - 1) / 2) return fail(ctx, CAC_EINVAL, "more edges requested than vertex pairs");
  if (m == 0) {
    if (draws_used) *draws_used = 0;
    return CAC_OK;
  }
  // The scan counts run in u32.
  if (m >= (1ull << 31)) return fail(ctx, CAC_EINVAL, "m must be < 2^31");
  hipStream_t s = ctx->stream;
  const uint64_t loop_key = (1ull << (2 * scale)) - 1;
  uint64_t draws = m + m / 4 + 4096;
  for (int attempt = 0; attempt < 8; ++attempt, draws = draws + draws / 2) {
    if (draws >= (1ull << 32)) return fail(ctx, CAC_EINVAL, "too many draws needed");
    uint64_t *keys = nullptr, *keys2 = nullptr, *cutoff = nullptr;
    uint32_t *vals = nullptr, *vals2 = nullptr, *partial = nullptr, *total = nullptr;
    uint8_t *mark = nullptr, *first = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    auto release = [&]() {
      (void)hipStreamSynchronize(s);
      for (void* p : {(void*)keys, (void*)keys2, (void*)vals, (void*)vals2, (void*)mark, (void*)first,
                      (void*)partial, (void*)total, (void*)cutoff, tmp})
        if (p) (void)hipFree(p);
    };
    hipError_t e = hipSuccess;
    if ((e = hipMalloc(&keys, draws * 8)) != hipSuccess || (e = hipMalloc(&keys2, draws * 8)) != hipSuccess ||
        (e = hipMalloc(&vals, draws * 4)) != hipSuccess || (e = hipMalloc(&vals2, draws * 4)) != hipSuccess ||
        (e = hipMalloc(&mark, draws)) != hipSuccess || (e = hipMalloc(&first, draws)) != hipSuccess ||
        (e = hipMalloc(&partial, (scan_tiles(draws) + 64) * 4)) != hipSuccess ||
        (e = hipMalloc(&total, 64)) != hipSuccess || (e = hipMalloc(&cutoff, 64)) != hipSuccess) {
      release();
      return hip_fail(ctx, e, "hipMalloc(R-MAT buffers)");
    }
    hipLaunchKernelGGL(k_rmat_keys, dim3(grid_for(draws, 256, 256 * 64)), dim3(256), 0, s, draws, scale, seed, ta, tab,
                       tabc, keys, vals);
    rocprim::double_buffer<uint64_t> kb(keys, keys2);
    rocprim::double_buffer<uint32_t> vb(vals, vals2);
    e = rocprim::radix_sort_pairs(nullptr, tmp_bytes, kb, vb, draws, 0, 2 * scale, s);
    if (e == hipSuccess) e = hipMalloc(&tmp, tmp_bytes);
    if (e == hipSuccess) e = rocprim::radix_sort_pairs(tmp, tmp_bytes, kb, vb, draws, 0, 2 * scale, s);
    if (e != hipSuccess) {
      release();
      return hip_fail(ctx, e, "R-MAT sort");
    }
    const uint64_t* sk = kb.current();
    const uint32_t* sv = vb.current();
    hipLaunchKernelGGL(k_rmat_first, dim3(grid_for(draws, 256, 256 * 64)), dim3(256), 0, s, sk, sv, draws, loop_key,
                       mark, first);
    hipLaunchKernelGGL(k_init_cutoff, dim3(1), dim3(1), 0, s, cutoff);
    e = device_exclusive_scan<uint32_t>(ByteLoad{mark}, CutoffStore{m, cutoff}, draws, partial, total, s);
    uint32_t distinct = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&distinct, total, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      release();
      return hip_fail(ctx, e, "R-MAT dedup");
    }
    if (distinct < m) {
      release();
      continue;  // not enough distinct pairs: draw more
    }
    e = device_exclusive_scan<uint32_t>(KeepLoad{first, sv, cutoff}, EmitStore{sk, scale, d_edges}, draws, partial,
                                        total, s);
    uint64_t cut = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&cut, cutoff, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    release();
    if (e != hipSuccess) return hip_fail(ctx, e, "R-MAT emit");
    if (draws_used) *draws_used = cut + 1;
    return CAC_OK;
  }
  return fail(ctx, CAC_EINVAL, "R-MAT: could not reach m distinct edges");
}

// ---- dense complete graph (C3 workload) and Erdos-Renyi edge lists ----------
namespace cac {
namespace {

// w(u, v) = 1 + floor(wmax * x / 2^32), x = word 0 of Philox(ctr = {pair_lo, pair_hi,
// 0, 0}, key = {seed, 'GENR'}), pair = u * n + v for u < v; symmetric, zero diagonal.
__global__ void __launch_bounds__(256) k_gen_complete(uint32_t n, uint32_t seed, uint32_t wmax, int64_t* A) {
  const uint64_t total = (uint64_t)n * n;
  for (uint64_t q = (uint64_t)blockIdx.x * 256 + threadIdx.x; q < total; q += (uint64_t)gridDim.x * 256) {
    const uint32_t r = (uint32_t)(q / n), c = (uint32_t)(q % n);
    if (r == c) {
      A[q] = 0;
      continue;
    }
    const uint32_t u = r < c ? r : c, v = r < c ? c : r;
    const uint64_t pair = (uint64_t)u * n + v;
    const u32x4 x = philox4x32((uint32_t)pair, (uint32_t)(pair >> 32), 0u, 0u, seed, kDomGen);
    A[q] = 1 + (int64_t)(((uint64_t)wmax * x.x) >> 32);
  }
}

}  // namespace
}  // namespace cac

extern "C" int cac_gen_complete_dev(cac_ctx* ctx, uint32_t n, uint32_t seed, uint32_t wmax, int64_t* d_adj) {
  using namespace cac;
  int rc = bind(ctx);
  if (rc) return rc;
  if (!d_adj || n < 2 || wmax < 1) return fail(ctx, CAC_EINVAL, "need n >= 2, wmax >= 1");
  hipLaunchKernelGGL(k_gen_complete, dim3(grid_for((uint64_t)n * n, 256, 256 * 64)), dim3(256), 0, ctx->stream, n, seed,
                     wmax, d_adj);
  CAC_LAUNCH_CHECK(ctx, "complete graph generator");
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CAC_OK;
}

extern "C" int cac_gen_er(uint32_t n, uint32_t threshold, uint32_t seed, uint32_t wmax, cac_edge* out, uint64_t cap,
                          uint64_t* m_out) {
  // pair (u < v) with index u*n+v is kept iff word 0 of Philox(pair; seed,'GENR') <
  // threshold; weight 1 + floor(wmax * word1 / 2^32).  Host-side (input generation).
  if (!m_out || n < 2) return CAC_EINVAL;
  uint64_t m = 0;
  for (uint32_t u = 0; u < n; ++u)
    for (uint32_t v = u + 1; v < n; ++v) {
      const uint64_t pair = (uint64_t)u * n + v;
      const cac::u32x4 x = cac::philox4x32((uint32_t)pair, (uint32_t)(pair >> 32), 0u, 0u, seed, cac::kDomGen);
      if (x.x >= threshold) continue;
      if (out && m < cap) out[m] = cac_edge{u, v, 1 + (((uint64_t)wmax * x.y) >> 32)};
      ++m;
    }
  *m_out = m;
  return (out && m > cap) ? CAC_ERANGE : CAC_OK;
}

extern "C" int cac_gen_ws(uint32_t n, uint32_t d, double beta, uint32_t seed, uint64_t weight, cac_edge* out,
                          uint64_t cap, uint64_t* m_out) {
  // ring lattice u -- u + j (j = 1..d), each edge rewired with probability beta
  // to (u, x), x uniform over the other n - 1 vertices; draws from
  // Philox(ctr = {u, j, 0, 0}, key = {seed, 'SWRL'}).  Host-side (input generation).
  constexpr uint32_t kDomWs = 0x5357524Cu;  // 'SWRL'
  if (!m_out || d == 0 || n <= 2 * (uint64_t)d || !(beta >= 0.0 && beta <= 1.0) || weight == 0) return CAC_EINVAL;
  const uint64_t m = (uint64_t)n * d;
  *m_out = m;
  if (!out) return CAC_OK;
  if (m > cap) return CAC_ERANGE;
  const uint64_t thr = (uint64_t)std::llround(beta * 4294967296.0);
  uint64_t i = 0;
  for (uint32_t u = 0; u < n; ++u)
    for (uint32_t j = 1; j <= d; ++j) {
      const cac::u32x4 x = cac::philox4x32(u, j, 0u, 0u, seed, kDomWs);
      uint32_t v = (uint32_t)(((uint64_t)u + j) % n);
      if ((uint64_t)x.x < thr) v = (uint32_t)(((uint64_t)u + 1 + (((uint64_t)(n - 1) * x.y) >> 32)) % n);
      out[i++] = cac_edge{u < v ? u : v, u < v ? v : u, weight};
    }
  return CAC_OK;
}
