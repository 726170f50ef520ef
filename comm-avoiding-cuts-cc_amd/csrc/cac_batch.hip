// cac_batch.hip -- batched square-root min cut over many independent graphs,
// and the planted-cut graph generator that feeds it.
//
// The reference solves one graph per MPI job (src/SquareRootCut.cpp:15-25).
// Here every (graph, trial) pair is one sampling job; jobs of different graphs
// share the sampling, densify and Karger-Stein launches (sqrt_cut_batch), so
// thousands of small graphs fill the GPU that one of them alone leaves idle.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <vector>

#include "cac_common.h"
#include "cac_internal.h"
#include "cac_ks.h"

namespace cac {
namespace {

constexpr uint32_t kDomPlanted = 0x504C4E54u;  // 'PLNT'
constexpr uint32_t kDomGnp = 0x474E5047u;      // 'GNPG'
constexpr uint32_t kMaxPlantedAttempts = 64;

uint32_t planted_n(uint32_t g, uint32_t seed, uint32_t nmin, uint32_t nmax) {
  const u32x4 r = philox4x32(g, 0xFFFFFFFFu, 0u, 0u, seed, kDomPlanted);
  const double u = ((double)r.x + 0.5) / 4294967296.0;
  const double lo = std::log((double)nmin), hi = std::log((double)nmax);
  uint32_t n = 2u * (uint32_t)(std::exp(lo + u * (hi - lo)) / 2.0);
  const uint32_t nlo = (nmin + 1) & ~1u, nhi = nmax & ~1u;
  return std::min(std::max(n, nlo), nhi);
}

uint32_t uf_find_host(std::vector<uint32_t>& par, uint32_t x) {
  while (par[x] != x) {
    par[x] = par[par[x]];
    x = par[x];
  }
  return x;
}

// One G(h, p) half on vertices [base, base + h): row u draws geometric skips
// from Philox(ctr = {g, 2 attempt + half, u, j}, key = {seed, 'PLNT'}) word 0
// and the weight 1 + floor(wmax * word1 / 2^32).
void planted_half(uint32_t g, uint32_t attempt, uint32_t half, uint32_t h, double p, uint32_t wmax, uint32_t seed,
                  uint32_t base, std::vector<cac_edge>& out) {
  const double l1mp = std::log1p(-p);
  for (uint32_t u = 0; u + 1 < h; ++u) {
    uint64_t v = u;
    for (uint32_t j = 0;; ++j) {
      const u32x4 r = philox4x32(g, 2 * attempt + half, u, j, seed, kDomPlanted);
      const double U = ((double)r.x + 0.5) / 4294967296.0;
      const double skip = std::floor(std::log(U) / l1mp);
      if (skip >= (double)h) break;
      v += 1 + (uint64_t)skip;
      if (v >= h) break;
      out.push_back(cac_edge{base + u, base + (uint32_t)v, 1 + (((uint64_t)wmax * r.y) >> 32)});
    }
  }
}

// Graph g: halves G(n/2, d/(n/2-1)) with weights U[1, wmax], joined by c
// distinct unit-weight cross edges, c uniform in [1, cmax].  Redrawn (attempt
// + 1) until the certificate holds: both halves connected and every weighted
// degree > c, so the c cross edges form the minimum cut among the cuts that
// separate the halves or isolate one vertex.  Returns false if no attempt passes.
bool planted_graph(uint32_t g, uint32_t seed, uint32_t n, double avg_degree, uint32_t wmax, uint32_t cmax,
                   std::vector<cac_edge>& out, uint32_t* c_out) {
  const uint32_t h = n / 2;
  const double p = std::min(1.0, avg_degree / (double)(h - 1));
  std::vector<uint32_t> par(n);
  std::vector<uint64_t> deg(n);
  for (uint32_t attempt = 0; attempt < kMaxPlantedAttempts; ++attempt) {
    out.clear();
    planted_half(g, attempt, 0, h, p, wmax, seed, 0, out);
    planted_half(g, attempt, 1, h, p, wmax, seed, h, out);
    const u32x4 rc = philox4x32(g, 0xFFFFFFFEu, attempt, 0u, seed, kDomPlanted);
    const uint32_t c = 1 + (uint32_t)(((uint64_t)cmax * rc.x) >> 32);
    const size_t cross0 = out.size();
    for (uint32_t j = 0; out.size() - cross0 < c; ++j) {
      const u32x4 r = philox4x32(g, 0xFFFFFFFDu, attempt, j, seed, kDomPlanted);
      const uint32_t x = (uint32_t)(((uint64_t)h * r.x) >> 32), y = h + (uint32_t)(((uint64_t)h * r.y) >> 32);
      bool dup = false;
      for (size_t k = cross0; k < out.size(); ++k) dup = dup || (out[k].u == x && out[k].v == y);
      if (!dup) out.push_back(cac_edge{x, y, 1});
    }
    for (uint32_t v = 0; v < n; ++v) par[v] = v;
    std::fill(deg.begin(), deg.end(), 0);
    for (size_t k = 0; k < out.size(); ++k) {
      deg[out[k].u] += out[k].w;
      deg[out[k].v] += out[k].w;
      if (k < cross0) {
        const uint32_t a = uf_find_host(par, out[k].u), b = uf_find_host(par, out[k].v);
        if (a != b) par[std::max(a, b)] = std::min(a, b);
      }
    }
    bool ok = true;
    for (uint32_t v = 0; v < n && ok; ++v) {
      const uint32_t r = uf_find_host(par, v);
      ok = deg[v] > c && r == (v < h ? 0u : h);
    }
    if (ok) {
      *c_out = c;
      return true;
    }
  }
  return false;
}

// Graph g of the hard family: each half is the union of k Hamiltonian cycles
// over its h = n/2 vertices.  Cycle i of half `half` visits the vertices in the
// order of a Fisher-Yates shuffle drawn from Philox(ctr = {g, 2^31 + 2 half,
// i, j}, key = {seed, 'PLNT'}) word 0, and all its edges weigh w_i = 1 +
// floor(wmax * word0 / 2^32) of Philox(ctr = {g, 2^31 + 2 half + 1, i, 0}).
// Every vertex of the half then has weighted degree delta = 2 sum w_i, and a
// cut inside the half crosses every cycle at least twice, so the half's
// minimum cut is delta.  c cross edges (c uniform in [1, cmax], distinct pairs)
// carry C = floor(cut_frac * min(delta_A, delta_B)) split as evenly as
// integers allow; with C < min delta the minimum cut of the graph is C,
// the cut between the halves, and it lies within (1 - cut_frac) of the
// minimum weighted degree.
bool planted_hard_graph(uint32_t g, uint32_t seed, uint32_t n, uint32_t k, uint32_t wmax, uint32_t cmax,
                        double cut_frac, std::vector<cac_edge>& out, uint32_t* C_out) {
  const uint32_t h = n / 2;
  if (h < 3 || k == 0) return false;
  out.clear();
  std::vector<uint32_t> perm(h);
  uint64_t delta[2] = {0, 0};
  for (uint32_t half = 0; half < 2; ++half) {
    const uint32_t base = half * h;
    for (uint32_t i = 0; i < k; ++i) {
      const u32x4 rw = philox4x32(g, 0x80000000u + 2 * half + 1, i, 0u, seed, kDomPlanted);
      const uint64_t w = 1 + (((uint64_t)wmax * rw.x) >> 32);
      delta[half] += 2 * w;
      for (uint32_t v = 0; v < h; ++v) perm[v] = v;
      for (uint32_t j = h - 1; j > 0; --j) {  // Fisher-Yates: position j takes one of [0, j]
        const u32x4 r = philox4x32(g, 0x80000000u + 2 * half, i, j, seed, kDomPlanted);
        std::swap(perm[j], perm[(uint32_t)(((uint64_t)(j + 1) * r.x) >> 32)]);
      }
      for (uint32_t j = 0; j < h; ++j) {
        const uint32_t a = base + perm[j], b = base + perm[(j + 1) % h];
        out.push_back(cac_edge{std::min(a, b), std::max(a, b), w});
      }
    }
  }
  const u32x4 rc = philox4x32(g, 0xFFFFFFFEu, 0x80000000u, 0u, seed, kDomPlanted);
  const uint32_t c = 1 + (uint32_t)(((uint64_t)cmax * rc.x) >> 32);
  const uint64_t C = (uint64_t)std::floor(cut_frac * (double)std::min(delta[0], delta[1]));
  if (C < c || C >= std::min(delta[0], delta[1])) return false;
  const size_t cross0 = out.size();
  for (uint32_t j = 0; out.size() - cross0 < c; ++j) {
    const u32x4 r = philox4x32(g, 0xFFFFFFFDu, 0x80000000u, j, seed, kDomPlanted);
    const uint32_t x = (uint32_t)(((uint64_t)h * r.x) >> 32), y = h + (uint32_t)(((uint64_t)h * r.y) >> 32);
    bool dup = false;
    for (size_t q = cross0; q < out.size(); ++q) dup = dup || (out[q].u == x && out[q].v == y);
    if (!dup) out.push_back(cac_edge{x, y, 0});
  }
  for (uint32_t j = 0; j < c; ++j) out[cross0 + j].w = C / c + (j < C % c ? 1 : 0);
  *C_out = (uint32_t)C;
  return true;
}

}  // namespace
}  // namespace cac

extern "C" {

int cac_gen_planted_hard(uint32_t first, uint32_t count, uint32_t seed, uint32_t nmin, uint32_t nmax,
                         uint32_t cycles, uint32_t wmax, uint32_t cmax, double cut_frac, cac_graph_desc* descs,
                         cac_edge* edges, uint64_t edge_cap, uint64_t* m_total, uint32_t* planted) {
  if (!descs || !edges || !m_total || !planted || nmin < 6 || nmax < nmin || cycles == 0 || wmax == 0 ||
      cmax == 0 || !(cut_frac > 0.0 && cut_frac < 1.0) || (uint64_t)wmax * cycles * 2 >= (1ull << 31))
    return CAC_EINVAL;
  std::vector<std::vector<cac_edge>> per(count);
  std::vector<uint32_t> ns(count), cs(count);
  std::atomic<uint32_t> next{0};
  std::atomic<int> failed{0};
  const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  auto work = [&]() {
    for (uint32_t i; (i = next.fetch_add(1)) < count;) {
      ns[i] = cac::planted_n(first + i, seed, nmin, nmax);
      if (!cac::planted_hard_graph(first + i, seed, ns[i], cycles, wmax, cmax, cut_frac, per[i], &cs[i])) failed = 1;
    }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < nth; ++t) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  if (failed) return CAC_EINVAL;
  uint64_t off = 0;
  for (uint32_t i = 0; i < count; ++i) {
    descs[i] = cac_graph_desc{off, (uint64_t)per[i].size(), ns[i], 0u};
    planted[i] = cs[i];
    off += per[i].size();
  }
  *m_total = off;
  if (off > edge_cap) return CAC_ERANGE;
  for (uint32_t i = 0; i < count; ++i) std::copy(per[i].begin(), per[i].end(), edges + descs[i].edge_offset);
  return CAC_OK;
}

int cac_gen_planted_sizes(uint32_t first, uint32_t count, uint32_t seed, uint32_t nmin, uint32_t nmax,
                          double avg_degree, uint32_t cmax, uint32_t* n_out, uint64_t* edge_bound) {
  if (!n_out || !edge_bound || nmin < 4 || nmax < nmin || !(avg_degree > 0)) return CAC_EINVAL;
  uint64_t bound = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const uint32_t n = cac::planted_n(first + i, seed, nmin, nmax);
    n_out[i] = n;
    // mean n d / 2 plus 8 standard deviations, plus the cross edges
    const double mean = (double)n * avg_degree / 2.0;
    bound += (uint64_t)(mean + 8.0 * std::sqrt(mean) + 64.0) + cmax;
  }
  *edge_bound = bound;
  return CAC_OK;
}

int cac_gen_planted(uint32_t first, uint32_t count, uint32_t seed, uint32_t nmin, uint32_t nmax, double avg_degree,
                    uint32_t wmax, uint32_t cmax, cac_graph_desc* descs, cac_edge* edges, uint64_t edge_cap,
                    uint64_t* m_total, uint32_t* planted) {
  if (!descs || !edges || !m_total || !planted || nmin < 4 || nmax < nmin || !(avg_degree > 0) || wmax == 0 ||
      cmax == 0)
    return CAC_EINVAL;
  std::vector<std::vector<cac_edge>> per(count);
  std::vector<uint32_t> ns(count), cs(count);
  std::atomic<uint32_t> next{0};
  std::atomic<int> failed{0};
  const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  auto work = [&]() {
    for (uint32_t i; (i = next.fetch_add(1)) < count;) {
      ns[i] = cac::planted_n(first + i, seed, nmin, nmax);
      if (!cac::planted_graph(first + i, seed, ns[i], avg_degree, wmax, cmax, per[i], &cs[i])) failed = 1;
    }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < nth; ++t) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  if (failed) return CAC_EINVAL;
  uint64_t off = 0;
  for (uint32_t i = 0; i < count; ++i) {
    descs[i] = cac_graph_desc{off, (uint64_t)per[i].size(), ns[i], 0u};
    planted[i] = cs[i];
    off += per[i].size();
  }
  *m_total = off;
  if (off > edge_cap) return CAC_ERANGE;
  for (uint32_t i = 0; i < count; ++i) std::copy(per[i].begin(), per[i].end(), edges + descs[i].edge_offset);
  return CAC_OK;
}

int cac_gen_gnp(uint32_t n, double avg_degree, uint32_t seed, uint32_t stream, uint32_t wmax, cac_edge* out,
                uint64_t cap, uint64_t* m_out) {
  if (!m_out || n < 2 || !(avg_degree > 0) || wmax == 0) return CAC_EINVAL;
  const double p = std::min(1.0, avg_degree / (double)(n - 1));
  const double l1mp = std::log1p(-p);
  uint64_t m = 0;
  for (uint32_t u = 0; u + 1 < n; ++u) {
    uint64_t v = u;
    for (uint32_t j = 0;; ++j) {
      const cac::u32x4 r = cac::philox4x32(stream, u, j, 0u, seed, cac::kDomGnp);
      double skip = 0.0;
      if (p < 1.0) {
        const double U = ((double)r.x + 0.5) / 4294967296.0;
        skip = std::floor(std::log(U) / l1mp);
        if (skip >= (double)n) break;
      }
      v += 1 + (uint64_t)skip;
      if (v >= n) break;
      if (out && m < cap) out[m] = cac_edge{u, (uint32_t)v, 1 + (((uint64_t)wmax * r.y) >> 32)};
      ++m;
    }
  }
  *m_out = m;
  return (out && m > cap) ? CAC_ERANGE : CAC_OK;
}

int cac_mincut_batch_dev(cac_ctx* ctx, const cac_edge* d_edges, const cac_graph_desc* descs, uint32_t count, double p,
                         uint32_t seed, int64_t* cuts, uint32_t* trials_out, void* stream) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if (!d_edges || !descs || !cuts || count == 0) return cac::fail(ctx, CAC_EINVAL, "null pointer or empty batch");
  if (!(p >= 0.0 && p < 1.0)) return cac::fail(ctx, CAC_EINVAL, "success probability must be in [0, 1)");
  std::vector<cac::SqrtJob> jobs;
  std::vector<uint64_t> first_slot(count + 1);
  for (uint32_t g = 0; g < count; ++g) {
    const cac_graph_desc& d = descs[g];
    if (d.n < 2 || d.m == 0) return cac::fail(ctx, CAC_EINVAL, "every graph needs n >= 2 and at least one edge");
    const uint32_t t = cac::sqrt_intermediate_size(d.n, d.m);
    const uint32_t T = d.trials ? d.trials : std::max<uint32_t>(1, cac::sqrt_trials_host(d.n, d.m, p));
    first_slot[g] = jobs.size();
    if (jobs.size() + T > 0xFFFFFFFFull) return cac::fail(ctx, CAC_EINVAL, "too many trials in one batch");
    for (uint32_t i = 0; i < T; ++i)
      jobs.push_back({d_edges + d.edge_offset, d.m, d.n, t, i, (uint32_t)jobs.size(), seed});
    if (trials_out) trials_out[g] = T;
  }
  first_slot[count] = jobs.size();
  hipStream_t s = cac::pick_stream(ctx, stream);
  int64_t* d_res = nullptr;
  CAC_HIP(ctx, hipMalloc((void**)&d_res, 8 * jobs.size()));
  rc = cac::sqrt_cut_batch(ctx, jobs, d_res, s, 0, true);
  if (rc == CAC_OK) {
    std::vector<int64_t> res(jobs.size());
    hipError_t e = hipMemcpyAsync(res.data(), d_res, 8 * res.size(), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = cac::hip_fail(ctx, e, "batch results");
    for (uint32_t g = 0; g < count && rc == CAC_OK; ++g)
      cuts[g] = *std::min_element(res.begin() + first_slot[g], res.begin() + first_slot[g + 1]);
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(d_res);
  return rc;
}

int cac_mincut_batch(cac_ctx* ctx, const cac_edge* edges, const cac_graph_desc* descs, uint32_t count, double p,
                     uint32_t seed, int64_t* cuts, uint32_t* trials_out) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if (!edges || !descs || count == 0) return cac::fail(ctx, CAC_EINVAL, "null pointer or empty batch");
  uint64_t total = 0;
  for (uint32_t g = 0; g < count; ++g) {
    const cac_graph_desc& d = descs[g];
    if (d.edge_offset > UINT64_MAX - d.m) return cac::fail(ctx, CAC_EINVAL, "edge range overflows");
    total = std::max(total, d.edge_offset + d.m);
    uint64_t wsum = 0;
    for (uint64_t i = 0; i < d.m; ++i) {
      const cac_edge& x = edges[d.edge_offset + i];
      if (x.u >= d.n || x.v >= d.n) return cac::fail(ctx, CAC_EINVAL, "edge endpoint >= its graph's n");
      if (x.w >= (1ull << 40)) return cac::fail(ctx, CAC_ERANGE, "edge weight >= 2^40");
      wsum += x.w;
      if (wsum >= (1ull << 61)) return cac::fail(ctx, CAC_ERANGE, "graph weight sum >= 2^61");
    }
  }
  hipStream_t s = ctx->stream;
  cac_edge* d = nullptr;
  CAC_HIP(ctx, hipMalloc((void**)&d, std::max<uint64_t>(total, 1) * sizeof(cac_edge)));
  hipError_t e = hipMemcpyAsync(d, edges, total * sizeof(cac_edge), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) rc = cac::hip_fail(ctx, e, "batch upload");
  if (rc == CAC_OK) rc = cac_mincut_batch_dev(ctx, d, descs, count, p, seed, cuts, trials_out, s);
  (void)hipStreamSynchronize(s);
  (void)hipFree(d);
  return rc;
}

}  // extern "C"
