// cac_api.hip -- min-cut entry points of the C ABI: trial-count rules, one
// sampling contraction (cac_sample_contract) and the low-concurrency sqrt-cut
// min cut (cac_mincut).
#include <algorithm>
#include <cmath>
#include <vector>

#include "cac_common.h"
#include "cac_internal.h"
#include "cac_ks.h"
#include "cac_scan.h"

namespace cac {
namespace {

// Flags bad ids (bit 0) and weights >= 2^40 (bit 1), and adds up the weights
// as two u64 counters, the low and the high 32 bits of every wavefront's sum
// (wsum[0], wsum[1]), so the total cannot wrap: with weights below 2^40 a
// wavefront's sum stays below 2^58 and the counters below 2^46 and 2^40 for
// fewer than 2^32 edges.
__global__ void k_check_edges(const cac_edge* e, uint64_t m, uint32_t n, uint32_t* bad, unsigned long long* wsum) {
  uint32_t b = 0;
  unsigned long long s = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
    const cac_edge x = e[i];
    if (x.u >= n || x.v >= n) b = 1;
    if (x.w >= (1ull << 40)) b |= 2;  // single weights this large would risk int64 sums
    else s += x.w;
  }
  if (b) atomicOr(bad, b);
  for (int d = 32; d > 0; d >>= 1) s += __shfl_xor(s, d, 64);
  if ((threadIdx.x & 63) == 0 && s) {
    atomicAdd(wsum, s & 0xFFFFFFFFull);
    atomicAdd(wsum + 1, s >> 32);
  }
}

__global__ void k_zero_words(uint32_t* p, int count) {
  if ((int)threadIdx.x < count) p[threadIdx.x] = 0;
}

// vmap[v] = rank of lab[v] among the canonical labels
struct LabRoot {
  const uint32_t* lab;
  __device__ uint32_t operator()(uint64_t i) const { return lab[i] == (uint32_t)i ? 1u : 0u; }
};
struct RankStore {
  uint32_t* rank;
  __device__ void operator()(uint64_t i, uint32_t ex, uint32_t f) const {
    if (f) rank[i] = ex;
  }
};
__global__ void k_vmap(const uint32_t* lab, const uint32_t* rank, uint32_t n, uint32_t* vmap) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) vmap[i] = rank[lab[i]];
}
// merged edges = upper-triangle non-zeros of the k x k matrix in row-major order
struct UpperNz {
  const int64_t* M;
  uint32_t k;
  __device__ uint64_t operator()(uint64_t i) const {
    const uint32_t r = (uint32_t)(i / k), c = (uint32_t)(i % k);
    return (c > r && M[i] != 0) ? 1u : 0u;
  }
};
struct EmitEdge {
  const int64_t* M;
  uint32_t k;
  cac_edge* out;
  uint64_t cap;
  __device__ void operator()(uint64_t i, uint64_t ex, uint64_t f) const {
    if (f && ex < cap) out[ex] = cac_edge{(uint32_t)(i / k), (uint32_t)(i % k), (uint64_t)M[i]};
  }
};

// Uploads a host edge list and validates it; *total = sum of weights.
int upload_edges(cac_ctx* ctx, const cac_edge* edges, uint64_t m, uint32_t n, cac_edge** d_out, hipStream_t s) {
  cac_edge* d = nullptr;
  CAC_HIP(ctx, hipMalloc((void**)&d, std::max<uint64_t>(m, 1) * sizeof(cac_edge) + 256));
  if (m) {
    hipError_t e = hipMemcpyAsync(d, edges, m * sizeof(cac_edge), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
      (void)hipFree(d);
      return hip_fail(ctx, e, "upload edges");
    }
  }
  uint32_t* flag = (uint32_t*)((char*)d + std::max<uint64_t>(m, 1) * sizeof(cac_edge));
  unsigned long long* wsum = (unsigned long long*)(flag + 2);
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, flag, 6);
  if (m) hipLaunchKernelGGL(k_check_edges, dim3(grid_for(m, 256, 4096)), dim3(256), 0, s, d, m, n, flag, wsum);
  uint32_t hb[6] = {0, 0, 0, 0, 0, 0};
  hipError_t e = hipMemcpyAsync(hb, flag, 24, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(d);
    return hip_fail(ctx, e, "edge check");
  }
  const unsigned long long lo = (unsigned long long)hb[2] | ((unsigned long long)hb[3] << 32);
  const unsigned long long hi = (unsigned long long)hb[4] | ((unsigned long long)hb[5] << 32);
  // total = hi * 2^32 + lo >= 2^61  <=>  hi + lo / 2^32 >= 2^29 (neither term wraps)
  const bool too_big = hi + (lo >> 32) >= (1ull << 29);
  if (hb[0] & 1u) {
    (void)hipFree(d);
    return fail(ctx, CAC_EINVAL, "edge endpoint >= n");
  }
  if ((hb[0] & 2u) || too_big) {
    (void)hipFree(d);
    return fail(ctx, CAC_ERANGE, "edge weights too large for int64 cut accumulation");
  }
  *d_out = d;
  return CAC_OK;
}

}  // namespace
}  // namespace cac

extern "C" {

uint32_t cac_ks_trials(uint32_t n, double p) { return cac::ks_trials_host(n, p); }
uint32_t cac_sqrt_trials(uint32_t n, uint64_t m, double p) { return cac::sqrt_trials_host(n, m, p); }
uint32_t cac_intermediate_size(uint32_t n, uint64_t m) { return cac::sqrt_intermediate_size(n, m); }
uint32_t cac_hc_draws(uint32_t n0) { return cac::hc_draws_host(n0); }

int cac_set_sampling_rule(cac_ctx* ctx, int rule) {
  if (!ctx) return CAC_EINVAL;
  if (rule != CAC_SAMPLE_LC && rule != CAC_SAMPLE_HC)
    return cac::fail(ctx, CAC_EINVAL, "sampling rule must be CAC_SAMPLE_LC or CAC_SAMPLE_HC");
  ctx->sample_rule = rule;
  return CAC_OK;
}

int cac_mincut_dev(cac_ctx* ctx, const cac_edge* d_edges, uint64_t m, uint32_t n, double p, uint32_t seed,
                   int64_t* cut, uint32_t* trials, void* stream) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if (!d_edges || !cut || m == 0 || n < 2) return cac::fail(ctx, CAC_EINVAL, "need n >= 2 and at least one edge");
  if (!(p >= 0.0 && p < 1.0)) return cac::fail(ctx, CAC_EINVAL, "success probability must be in [0, 1)");
  hipStream_t s = cac::pick_stream(ctx, stream);
  const uint32_t t = cac::sqrt_intermediate_size(n, m);
  const uint32_t T = std::max<uint32_t>(1, cac::sqrt_trials_host(n, m, p));
  std::vector<uint32_t> ids(T);
  for (uint32_t i = 0; i < T; ++i) ids[i] = i;
  int64_t* d_res = nullptr;
  CAC_HIP(ctx, hipMalloc((void**)&d_res, 8 * (size_t)T));
  rc = cac::sqrt_cut_run(ctx, d_edges, m, n, t, seed, ids, d_res, s);
  if (rc == CAC_OK) {
    std::vector<int64_t> res(T);
    hipError_t e = hipMemcpyAsync(res.data(), d_res, 8 * (size_t)T, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = cac::hip_fail(ctx, e, "mincut results");
    else {
      *cut = *std::min_element(res.begin(), res.end());
      if (trials) *trials = T;
    }
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(d_res);
  return rc;
}

int cac_sqrt_cut_trials(cac_ctx* ctx, const cac_edge* edges, uint64_t m, uint32_t n, uint32_t t, uint32_t seed,
                        uint32_t first_trial, uint32_t ntrials, int64_t* raw_cuts) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if (!edges || !raw_cuts || m == 0 || n < 2 || ntrials == 0 || t < 2)
    return cac::fail(ctx, CAC_EINVAL, "need n >= 2, t >= 2, ntrials >= 1 and at least one edge");
  if (t > n) t = n;
  hipStream_t s = ctx->stream;
  cac_edge* d = nullptr;
  rc = cac::upload_edges(ctx, edges, m, n, &d, s);
  if (rc) return rc;
  std::vector<uint32_t> ids(ntrials);
  for (uint32_t i = 0; i < ntrials; ++i) ids[i] = first_trial + i;
  int64_t* d_res = nullptr;
  hipError_t e = hipMalloc((void**)&d_res, 8 * (size_t)ntrials);
  if (e != hipSuccess) rc = cac::hip_fail(ctx, e, "results");
  if (rc == CAC_OK) rc = cac::sqrt_cut_run(ctx, d, m, n, t, seed, ids, d_res, s);
  if (rc == CAC_OK) {
    e = hipMemcpyAsync(raw_cuts, d_res, 8 * (size_t)ntrials, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = cac::hip_fail(ctx, e, "results");
  }
  (void)hipStreamSynchronize(s);
  if (d_res) (void)hipFree(d_res);
  (void)hipFree(d);
  return rc;
}

int cac_mincut(cac_ctx* ctx, const cac_edge* edges, uint64_t m, uint32_t n, double p, uint32_t seed, int64_t* cut,
               uint32_t* trials) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if (!edges || !cut || m == 0 || n < 2) return cac::fail(ctx, CAC_EINVAL, "need n >= 2 and at least one edge");
  cac_edge* d = nullptr;
  rc = cac::upload_edges(ctx, edges, m, n, &d, ctx->stream);
  if (rc) return rc;
  rc = cac_mincut_dev(ctx, d, m, n, p, seed, cut, trials, ctx->stream);
  (void)hipFree(d);
  return rc;
}

int cac_sample_contract(cac_ctx* ctx, const cac_edge* edges, uint64_t m, uint32_t n, uint32_t t, uint32_t seed,
                        uint32_t trial, uint32_t* vmap_out, uint32_t* n_out, cac_edge* merged_out, uint64_t merged_cap,
                        uint64_t* m_out) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if (!edges || !vmap_out || !n_out || !m_out || m == 0 || n < 2 || t < 2)
    return cac::fail(ctx, CAC_EINVAL, "need n >= 2, t >= 2 and at least one edge");
  if (t > n) t = n;
  hipStream_t s = ctx->stream;
  cac_edge* d_e = nullptr;
  rc = cac::upload_edges(ctx, edges, m, n, &d_e, s);
  if (rc) return rc;
  cac::SampleOut so;
  rc = cac::sample_contract_one(ctx, d_e, m, n, t, seed, trial, &so, s);
  if (rc == CAC_OK) {
    const uint32_t k = so.k;
    *n_out = k;
    // vmap and merged edges, both extracted on the device
    size_t bytes = cac::carve_size((size_t)n * 4) * 2 + cac::carve_size((cac::scan_tiles((uint64_t)n) + 8) * 4) +
                   cac::carve_size((cac::scan_tiles((uint64_t)k * k) + 8) * 8) +
                   cac::carve_size(std::max<uint64_t>(merged_cap, 1) * sizeof(cac_edge)) + 256;
    void* arena = nullptr;
    hipError_t e = hipMalloc(&arena, bytes);
    if (e != hipSuccess) rc = cac::hip_fail(ctx, e, "sample_contract outputs");
    else {
      cac::Carver cv(arena);
      uint32_t* rank = cv.take<uint32_t>(n);
      uint32_t* vmap = cv.take<uint32_t>(n);
      uint32_t* part32 = cv.take<uint32_t>(cac::scan_tiles((uint64_t)n) + 8);
      uint64_t* part64 = cv.take<uint64_t>(cac::scan_tiles((uint64_t)k * k) + 8);
      cac_edge* d_out = cv.take<cac_edge>(std::max<uint64_t>(merged_cap, 1));
      uint64_t* d_cnt = part64 + cac::scan_tiles((uint64_t)k * k) + 2;
      e = cac::device_exclusive_scan<uint32_t>(cac::LabRoot{so.lab}, cac::RankStore{rank}, n, part32,
                                               (uint32_t*)nullptr, s);
      if (e == hipSuccess) {
        hipLaunchKernelGGL(cac::k_vmap, dim3(cac::grid_for(n, 256, 4096)), dim3(256), 0, s, (const uint32_t*)so.lab,
                           (const uint32_t*)rank, n, vmap);
        e = cac::device_exclusive_scan<uint64_t>(cac::UpperNz{so.mat, k}, cac::EmitEdge{so.mat, k, d_out, merged_cap},
                                                 (uint64_t)k * k, part64, d_cnt, s);
      }
      uint64_t cnt = 0;
      if (e == hipSuccess) e = hipMemcpyAsync(vmap_out, vmap, (size_t)n * 4, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipMemcpyAsync(&cnt, d_cnt, 8, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e == hipSuccess && merged_out && cnt)
        e = hipMemcpy(merged_out, d_out, std::min(cnt, merged_cap) * sizeof(cac_edge), hipMemcpyDeviceToHost);
      if (e != hipSuccess) rc = cac::hip_fail(ctx, e, "sample_contract outputs");
      else {
        *m_out = cnt;
        if (cnt > merged_cap) rc = cac::fail(ctx, CAC_ERANGE, "merged_cap too small for the merged edge list");
      }
      (void)hipStreamSynchronize(s);
      (void)hipFree(arena);
    }
  }
  cac::sample_out_free(&so);
  (void)hipFree(d_e);
  return rc;
}

}  // extern "C"

namespace cac {
int approx_cut_run(cac_ctx* ctx, const cac_edge* d_edges, uint64_t m, uint32_t n, uint32_t copies, uint32_t iters,
                   uint32_t seed, uint32_t* first_disconnected, hipStream_t s);
}

extern "C" {

uint32_t cac_approx_trials(double p) {
  double alpha = std::log(1 - std::exp(-(double)0.75));
  return (uint32_t)std::ceil(std::log((double)1.0 - p) / alpha);
}

int cac_approx_cut(cac_ctx* ctx, const cac_edge* edges, uint64_t m, uint32_t n, double p, uint32_t seed,
                   uint64_t* approx, uint32_t* trials) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if (!edges || !approx || m == 0 || n < 2) return cac::fail(ctx, CAC_EINVAL, "need n >= 2 and at least one edge");
  if (!(p > 0.0 && p < 1.0)) return cac::fail(ctx, CAC_EINVAL, "success probability must be in (0, 1)");
  hipStream_t s = ctx->stream;
  long double total = 0;
  for (uint64_t e = 0; e < m; ++e) total += edges[e].w;
  if (total < 1) return cac::fail(ctx, CAC_EINVAL, "total weight must be positive");
  const uint32_t iters = (uint32_t)std::ceil(std::log((double)total));
  const uint32_t T = cac_approx_trials(p);
  cac_edge* d = nullptr;
  rc = cac::upload_edges(ctx, edges, m, n, &d, s);
  if (rc) return rc;
  uint32_t first = 0;
  rc = cac::approx_cut_run(ctx, d, m, n, T, iters, seed, &first, s);
  (void)hipFree(d);
  if (rc) return rc;
  *approx = 1ull << first;
  if (trials) *trials = T * first;
  return CAC_OK;
}

}  // extern "C"
