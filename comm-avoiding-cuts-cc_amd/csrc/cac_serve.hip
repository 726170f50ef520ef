// cac_serve.hip -- a serving loop for connected-components and min-cut queries
// on one GPU (no counterpart in the reference, whose executables solve one
// graph per MPI job).
//
// Design: W worker threads, each with its own context (HIP stream + scratch).
// A free worker takes the oldest pending query, then every other pending query
// of the same kind that still fits the HBM budget, up to max_batch, and runs
// them as one launch group:
//   - CC: the graphs are laid side by side (vertex ids offset) and one CC
//     pipeline labels all of them; labels are canonical minimum ids, so each
//     query's labels and component count are exactly its alone result;
//   - min cut: every (query, trial) job of the group shares the sampling,
//     densify and Karger-Stein launches (sqrt_cut_batch); each job carries its
//     query's seed, so queries of any seeds share launches.
// Groups on different workers overlap on the device, so a min-cut group takes
// at most an equal share (1 / workers) of the pending min-cut work (trials x
// edges), the oldest query always: under a backlog every worker gets a share,
// and one group's latency-bound stretches (its Karger-Stein pick chains) run
// beside the others' sampling instead of one after another in a single group.
// A query is admitted only
// when its footprint estimate fits beside the groups in flight, so requests
// that arrive while others run are admitted as memory and workers free up.  A
// min-cut group runs its trials in chunks, so it is charged its uploads plus
// at most an equal per-worker share of the budget for scratch (or its largest
// query's scratch, if more); that share is the chunk budget of its batch.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "cac_common.h"
#include "cac_internal.h"
#include "cac_ks.h"
#include "cac_trial.h"

namespace cac {

size_t cc_tmp_words(uint64_t m, uint32_t n, int label_mode);
int cc_launch(cac_ctx* ctx, const cac_uedge* d_edges, uint64_t m, uint32_t n, int label_mode, uint32_t* d_lab,
              uint32_t* d_tmp, hipStream_t s);

namespace {

constexpr uint32_t kServeMaxT = 4096;              // dense base size cap for explicit trial budgets
constexpr uint64_t kServeMaxGroupEdges = 1ull << 27;  // edges per CC launch group
// Trials of one query whose sampling and Karger-Stein scratch count towards its
// admitted footprint: the group's trials run in chunks of at most that many
// per query (two waves of sampling workgroups on 256 CUs).
constexpr uint32_t kServeChunkTrials = 1024;

struct Pending {
  uint64_t id;
  cac_query q;
  double submit_s;
  size_t bytes;             // upload + scratch
  size_t upload;            // its edges, when it brings them
  size_t scratch;           // device scratch of its work (for a min cut: its trials' chunk)
  uint64_t work;            // a min cut's trials x edges (its share of a group), 0 for CC
  const cac_edge* d_edges;  // a registered graph's device edges (checked once), or null
};

// A graph registered with the server: its edges stay in device memory, checked
// once, and queries name it instead of bringing their edges.
struct Resident {
  cac_edge* d_edges;
  uint64_t m;
  uint32_t n;
  size_t bytes;
  uint32_t users;  // queries submitted and not yet finished
  bool dropped;    // unregistered: freed when the last user finishes
};

// trials == 0: the reference's rule, as cac_mincut runs it (SquareRootCut::
// intermediate_size and numberOfTrials, src/SquareRootCut.cpp:63-83);
// trials > 0: that many trials on a contracted graph of at most kServeMaxT
// vertices (the dense engine's fast sizes).
void mincut_plan(uint32_t n, uint64_t m, double p, uint32_t trials, uint32_t* t, uint32_t* T) {
  if (trials == 0) {
    *t = sqrt_intermediate_size(n, m);
    *T = std::max<uint32_t>(1, sqrt_trials_host(n, m, p));
  } else {
    *t = std::min<uint32_t>(sqrt_intermediate_size(n, m), kServeMaxT);
    *T = trials;
  }
}

// Device bytes a query holds while it runs: its upload (a registered graph's
// edges are already resident and not counted again) and its scratch.  A min
// cut's scratch covers a chunk of up to kServeChunkTrials trials, with the
// wide form's buffers for edge lists of 2^20 edges or more.
void footprint(const cac_query& q, bool resident, size_t* upload, size_t* scratch) {
  *upload = resident ? 0 : carve_size(q.m * sizeof(cac_edge));
  if (q.kind == CAC_QUERY_CC) {
    *scratch = carve_size(q.m * sizeof(cac_uedge)) + carve_size((size_t)q.n * 4) + cc_tmp_words(q.m, q.n, 0) * 4 + 4096;
    return;
  }
  uint32_t t, T;
  mincut_plan(q.n, q.m, q.p, q.trials, &t, &T);
  SqrtJob j{nullptr, q.m, q.n, t, 0, 0, 0};
  const size_t wide = q.m >= (1ull << 20) ? wide_trial_bytes(q.m, q.n) : 0;
  *scratch = (size_t)std::min(T, kServeChunkTrials) * (job_sample_bytes(j) + job_ks_bytes(j) + wide) + 4096;
}

// Bytes a launch group holds at once.  CC graphs are laid side by side, so a
// CC group holds all of its scratch.  A min-cut group runs its trials in
// chunks sized to the bytes it was admitted with, so it holds at most
// max(cap, its largest query's scratch) beside its uploads.
size_t group_bytes(int kind, size_t upload, size_t scratch, size_t scratch_max, size_t cap) {
  if (kind == CAC_QUERY_CC) return upload + scratch;
  return upload + std::min(scratch, std::max(cap, scratch_max));
}

}  // namespace
}  // namespace cac

struct cac_server {
  int device = 0;
  uint64_t budget = 0;
  uint32_t max_batch = 1;
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<cac::Pending> pending;
  std::vector<cac_query_result> done;
  uint64_t inflight = 0, next_id = 0;
  uint64_t pending_work = 0;  // sum of the pending min-cut queries' work
  uint64_t peak_inflight = 0, groups = 0;
  std::unordered_map<uint64_t, cac::Resident> graphs;  // registered graphs by handle
  uint64_t next_graph = 1, resident = 0;               // resident: bytes of registered graphs
  bool stop = false;
  std::vector<std::thread> workers;
  std::vector<cac_ctx*> ctxs;

  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace cac {
namespace {

cac_query_result blank(const Pending& p) {
  cac_query_result r{};
  r.id = p.id;
  r.status = CAC_OK;
  r.submit_s = p.submit_s;
  return r;
}

// Cheap checks at submit time; edge ids and weights are checked on the device
// (a registered graph's when it was registered).
int validate(const cac_query& q, bool resident) {
  if (q.n < 1 || (q.m && !q.edges && !resident)) return CAC_EINVAL;
  if (q.kind == CAC_QUERY_MINCUT && (q.n < 2 || q.m == 0 || !(q.p >= 0.0 && q.p < 1.0))) return CAC_EINVAL;
  return CAC_OK;
}

// CC group staging: query edges (cac_edge) -> side-by-side cac_uedge with the
// query's vertex offset.  An edge with an id >= n becomes the loop (off, off)
// and flags the query, so no kernel of the group can index outside it.
__global__ void __launch_bounds__(256) k_serve_prep_cc(const cac_edge* __restrict__ in, uint64_t m, uint32_t n,
                                                       uint32_t off, cac_uedge* __restrict__ out, uint32_t* bad) {
  uint32_t b = 0;
  for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < m; e += (uint64_t)gridDim.x * 256) {
    const cac_edge x = in[e];
    const bool ok = x.u < n && x.v < n;
    out[e] = ok ? cac_uedge{x.u + off, x.v + off} : cac_uedge{off, off};
    b |= ok ? 0u : 1u;
  }
  if (b) atomicOr(bad, 1u);
}

// Min-cut group staging: flags ids >= n (chk[0] bit 0) and weights >= 2^40
// (bit 1); such an edge becomes a weight-0 loop.  Sums the weights, plus the
// weights' high words so a wrapped 64-bit sum is still caught (chk[1], chk[2]).
__global__ void __launch_bounds__(256) k_serve_check_w(cac_edge* __restrict__ e, uint64_t m, uint32_t n,
                                                       unsigned long long* chk) {
  __shared__ unsigned long long red[2][4];
  unsigned long long s = 0, hs = 0;
  uint32_t b = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (uint64_t)gridDim.x * 256) {
    const cac_edge x = e[i];
    const uint32_t f = (x.u >= n || x.v >= n ? 1u : 0u) | (x.w >= (1ull << 40) ? 2u : 0u);
    if (f) {
      e[i] = cac_edge{0, 0, 0};
      b |= f;
    } else {
      s += x.w;
      hs += x.w >> 32;
    }
  }
  for (int d = 32; d > 0; d >>= 1) {
    s += __shfl_down(s, d, 64);
    hs += __shfl_down(hs, d, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = s;
    red[1][threadIdx.x >> 6] = hs;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      s += red[0][w];
      hs += red[1][w];
    }
    if (s) atomicAdd(&chk[1], s);
    if (hs) atomicAdd(&chk[2], hs);
  }
  if (b) atomicOr(&chk[0], (unsigned long long)b);
}

__global__ void k_zero_u64(unsigned long long* p, uint32_t count) {
  for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) p[i] = 0;
}
__global__ void k_zero_u32s(uint32_t* p, uint32_t count) {
  for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) p[i] = 0;
}

void run_cc_group(cac_ctx* ctx, std::vector<Pending>& grp, std::vector<cac_query_result>& res) {
  uint64_t M = 0, N = 0;
  std::vector<uint64_t> voff(grp.size()), eoff(grp.size());
  for (size_t i = 0; i < grp.size(); ++i) {
    voff[i] = N;
    eoff[i] = M;
    N += grp[i].q.n;
    M += grp[i].q.m;
  }
  const size_t tw = cc_tmp_words(M, (uint32_t)N, CAC_LABEL_MIN_ID);
  // the worker's grow-only buffer: no allocation (and no device-wide wait
  // of a free) per group once it has grown
  int rc = grow_buffer(ctx, &ctx->serve_buf, &ctx->serve_bytes,
                       carve_size(M * sizeof(cac_edge)) + carve_size(M * sizeof(cac_uedge)) + carve_size(N * 4) +
                           carve_size(tw * 4) + carve_size(grp.size() * 4) + 2048);
  BufPin pin(ctx, kBufServe);
  hipError_t e = hipSuccess;
  std::vector<uint32_t> lab(N), bad(grp.size(), 0);
  hipStream_t s = ctx->stream;
  if (rc == CAC_OK) {
    Carver cv(ctx->serve_buf);
    cac_edge* d_in = cv.take<cac_edge>(M);
    cac_uedge* d_e = cv.take<cac_uedge>(M);
    uint32_t* d_l = cv.take<uint32_t>(N);
    uint32_t* d_t = cv.take<uint32_t>(tw);
    uint32_t* d_bad = cv.take<uint32_t>(grp.size());
    hipLaunchKernelGGL(k_zero_u32s, dim3(1), dim3(256), 0, s, d_bad, (uint32_t)grp.size());
    for (size_t i = 0; i < grp.size() && e == hipSuccess; ++i) {
      const cac_query& q = grp[i].q;
      if (!q.m) continue;
      const cac_edge* src = grp[i].d_edges;  // a registered graph needs no upload
      if (!src) {
        e = hipMemcpyAsync(d_in + eoff[i], q.edges, q.m * sizeof(cac_edge), hipMemcpyHostToDevice, s);
        src = d_in + eoff[i];
      }
      hipLaunchKernelGGL(k_serve_prep_cc, dim3(grid_for(q.m, 256, 2048)), dim3(256), 0, s, src, q.m, q.n,
                         (uint32_t)voff[i], d_e + eoff[i], d_bad + i);
    }
    if (e != hipSuccess) rc = hip_fail(ctx, e, "serve cc upload");
    if (rc == CAC_OK) rc = cc_launch(ctx, d_e, M, (uint32_t)N, CAC_LABEL_MIN_ID, d_l, d_t, s);
    if (rc == CAC_OK) {
      e = hipMemcpyAsync(lab.data(), d_l, N * 4, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipMemcpyAsync(bad.data(), d_bad, grp.size() * 4, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) rc = hip_fail(ctx, e, "serve cc results");
    }
  }
  (void)hipStreamSynchronize(s);
  for (size_t i = 0; i < grp.size(); ++i) {
    cac_query_result& r = res[i];
    r.status = rc != CAC_OK ? rc : (bad[i] ? CAC_EINVAL : CAC_OK);
    if (r.status != CAC_OK) continue;
    uint32_t nc = 0;
    const uint32_t off = (uint32_t)voff[i];
    for (uint32_t v = 0; v < grp[i].q.n; ++v) {
      const uint32_t l = lab[off + v] - off;  // canonical ids stay inside the query's range
      nc += (l == v);
      if (grp[i].q.labels_out) grp[i].q.labels_out[v] = l;
    }
    r.ncomp = nc;
  }
}

void run_mincut_group(cac_ctx* ctx, std::vector<Pending>& grp, std::vector<cac_query_result>& res, size_t budget) {
  uint64_t M = 0;  // edges to upload: registered graphs are already on the device
  size_t jobs_max = 0;
  for (auto& p : grp) {
    M += p.d_edges ? 0 : p.q.m;
    uint32_t t, T;
    mincut_plan(p.q.n, p.q.m, p.q.p, p.q.trials, &t, &T);
    jobs_max += T;
  }
  // the worker's grow-only buffer: no allocation (and no device-wide wait of
  // a free) per group once it has grown
  int rc = grow_buffer(ctx, &ctx->serve_buf, &ctx->serve_bytes,
                       carve_size(std::max<uint64_t>(M, 1) * sizeof(cac_edge)) + carve_size(24 * grp.size()) +
                           carve_size(8 * std::max<size_t>(jobs_max, 1)) + 1024);
  BufPin pin(ctx, kBufServe);
  hipStream_t s = ctx->stream;
  cac_edge* d_e = nullptr;
  int64_t* d_res = nullptr;
  unsigned long long* d_chk = nullptr;
  hipError_t e = hipSuccess;
  if (rc == CAC_OK) {
    Carver cv(ctx->serve_buf);
    d_e = cv.take<cac_edge>(std::max<uint64_t>(M, 1));
    d_chk = cv.take<unsigned long long>(3 * grp.size());
    d_res = cv.take<int64_t>(std::max<size_t>(jobs_max, 1));
  }
  std::vector<uint64_t> eoff(grp.size());
  std::vector<unsigned long long> chk(3 * grp.size());
  if (rc == CAC_OK) {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(256), 0, s, d_chk, (uint32_t)(3 * grp.size()));
    uint64_t off = 0;
    for (size_t i = 0; i < grp.size() && e == hipSuccess; ++i) {
      const cac_query& q = grp[i].q;
      eoff[i] = off;
      if (grp[i].d_edges) continue;  // checked when it was registered
      e = hipMemcpyAsync(d_e + off, q.edges, q.m * sizeof(cac_edge), hipMemcpyHostToDevice, s);
      hipLaunchKernelGGL(k_serve_check_w, dim3(grid_for(q.m, 256, 2048)), dim3(256), 0, s, d_e + off, q.m, q.n,
                         d_chk + 3 * i);
      off += q.m;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(chk.data(), d_chk, 24 * grp.size(), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);  // bad queries leave the group before any sampling
    if (e != hipSuccess) rc = hip_fail(ctx, e, "serve mincut upload");
  }
  std::vector<SqrtJob> jobs;
  std::vector<size_t> first(grp.size() + 1, 0);
  for (size_t i = 0; i < grp.size() && rc == CAC_OK; ++i) {
    const cac_query& q = grp[i].q;
    first[i] = jobs.size();
    if (chk[3 * i] & 1) {
      res[i].status = CAC_EINVAL;
      continue;
    }
    if ((chk[3 * i] & 2) || chk[3 * i + 1] >= (1ull << 61) || chk[3 * i + 2] >= (1ull << 29)) {
      res[i].status = CAC_ERANGE;
      continue;
    }
    uint32_t t, T;
    mincut_plan(q.n, q.m, q.p, q.trials, &t, &T);
    res[i].t = t;
    res[i].trials = T;
    const cac_edge* qe = grp[i].d_edges ? grp[i].d_edges : d_e + eoff[i];
    for (uint32_t k = 0; k < T; ++k) jobs.push_back({qe, q.m, q.n, t, k, (uint32_t)jobs.size(), q.seed});
  }
  first[grp.size()] = jobs.size();
  if (rc == CAC_OK && !jobs.empty()) {
    // the largest graphs' trials first: they are the launch's longest
    // workgroups, so they start while the small ones fill the other CUs
    // (results go by slot, so the order changes no cut)
    std::stable_sort(jobs.begin(), jobs.end(), [](const SqrtJob& x, const SqrtJob& y) { return x.m > y.m; });
    if (rc == CAC_OK) rc = sqrt_cut_batch(ctx, jobs, d_res, s, budget);
  }
  std::vector<int64_t> raw(jobs.size());
  if (rc == CAC_OK && !jobs.empty()) {
    e = hipMemcpyAsync(raw.data(), d_res, 8 * raw.size(), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(ctx, e, "serve mincut results");
  }
  (void)hipStreamSynchronize(s);
  for (size_t i = 0; i < grp.size(); ++i) {
    if (rc != CAC_OK) res[i].status = rc;
    if (res[i].status == CAC_OK)
      res[i].cut = *std::min_element(raw.begin() + first[i], raw.begin() + first[i + 1]);
  }
}

void worker_loop(cac_server* S, cac_ctx* ctx) {
  for (;;) {
    std::vector<Pending> grp;
    size_t bytes = 0, chunk_budget = 0;
    double admit = 0;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      for (;;) {
        if (S->stop) return;
        if (!S->pending.empty() && S->inflight + S->resident + S->pending.front().bytes <= S->budget) break;
        S->cv.wait(lk);
      }
      Pending head = S->pending.front();
      S->pending.pop_front();
      // a min-cut group's share of the pending work (the oldest query always)
      const uint64_t share = S->pending_work / std::max<size_t>(S->ctxs.size(), 1);
      uint64_t work = head.work;
      S->pending_work -= head.work;
      // a min-cut group's scratch is capped at an equal share of the budget
      // per worker: its trials then run in chunks of that size
      const size_t cap = S->budget / std::max<size_t>(S->ctxs.size(), 1);
      size_t upload = head.upload, scratch = head.scratch, scratch_max = head.scratch;
      bytes = group_bytes(head.q.kind, upload, scratch, scratch_max, cap);
      uint64_t edges = head.q.m;
      grp.push_back(head);
      for (auto it = S->pending.begin(); it != S->pending.end() && grp.size() < S->max_batch;) {
        const bool same = it->q.kind == head.q.kind;
        const size_t with = group_bytes(head.q.kind, upload + it->upload, scratch + it->scratch,
                                        std::max(scratch_max, it->scratch), cap);
        const bool fits = S->inflight + S->resident + with <= S->budget &&
                          (head.q.kind != CAC_QUERY_CC || edges + it->q.m <= kServeMaxGroupEdges);
        const bool ids_fit = head.q.kind != CAC_QUERY_CC || [&] {
          uint64_t nsum = 0;
          for (auto& g : grp) nsum += g.q.n;
          return nsum + it->q.n < (1ull << 32);
        }();
        const bool in_share = head.q.kind == CAC_QUERY_CC || work + it->work <= share;
        if (same && fits && ids_fit && in_share) {
          work += it->work;
          S->pending_work -= it->work;
          upload += it->upload;
          scratch += it->scratch;
          scratch_max = std::max(scratch_max, it->scratch);
          bytes = with;
          edges += it->q.m;
          grp.push_back(*it);
          it = S->pending.erase(it);
        } else {
          ++it;
        }
      }
      S->inflight += bytes;
      S->peak_inflight = std::max(S->peak_inflight, S->inflight);
      ++S->groups;
      admit = S->now();
      chunk_budget = bytes - upload;
    }
    std::vector<cac_query_result> res(grp.size());
    for (size_t i = 0; i < grp.size(); ++i) {
      res[i] = blank(grp[i]);
      res[i].admit_s = admit;
      res[i].batch_size = (uint32_t)grp.size();
    }
    (void)bind(ctx);
    if (grp[0].q.kind == CAC_QUERY_CC)
      run_cc_group(ctx, grp, res);
    else
      run_mincut_group(ctx, grp, res, chunk_budget);
    const double fin = S->now();
    std::vector<cac_edge*> freed;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->inflight -= bytes;
      for (auto& r : res) {
        r.done_s = fin;
        S->done.push_back(r);
      }
      for (auto& p : grp) {  // registered graphs: one user fewer; free unregistered ones now unused
        if (!p.q.graph) continue;
        auto it = S->graphs.find(p.q.graph);
        if (it == S->graphs.end() || --it->second.users > 0 || !it->second.dropped) continue;
        freed.push_back(it->second.d_edges);
        S->resident -= it->second.bytes;
        S->graphs.erase(it);
      }
    }
    for (cac_edge* d : freed) (void)hipFree(d);
    S->cv.notify_all();
  }
}

}  // namespace
}  // namespace cac

extern "C" {

int cac_serve_mincut_plan(uint32_t n, uint64_t m, double p, uint32_t trials_req, uint32_t* t, uint32_t* trials) {
  if (!t || !trials || n < 2 || m == 0) return CAC_EINVAL;
  cac::mincut_plan(n, m, p, trials_req, t, trials);
  return CAC_OK;
}

int cac_server_create(int device, uint32_t workers, uint64_t hbm_budget, uint32_t max_batch, cac_server** out) {
  if (!out || workers == 0 || workers > 16 || hbm_budget == 0 || max_batch == 0) return CAC_EINVAL;
  *out = nullptr;
  cac_server* S = new (std::nothrow) cac_server();
  if (!S) return CAC_ENOMEM;
  S->device = device;
  S->budget = hbm_budget;
  S->max_batch = max_batch;
  S->t0 = std::chrono::steady_clock::now();
  for (uint32_t w = 0; w < workers; ++w) {
    cac_ctx* c = nullptr;
    int rc = cac_ctx_create(device, &c);
    if (rc != CAC_OK) {
      for (auto x : S->ctxs) cac_ctx_destroy(x);
      delete S;
      return rc;
    }
    c->grow_wide = true;  // groups differ from call to call: regrow with room to spare
    S->ctxs.push_back(c);
  }
  for (uint32_t w = 0; w < workers; ++w) S->workers.emplace_back(cac::worker_loop, S, S->ctxs[w]);
  *out = S;
  return CAC_OK;
}

double cac_server_clock(cac_server* S) { return S ? S->now() : 0.0; }

int cac_server_stats(cac_server* S, uint64_t* peak_inflight_bytes, uint64_t* groups) {
  if (!S || !peak_inflight_bytes || !groups) return CAC_EINVAL;
  std::lock_guard<std::mutex> lk(S->mu);
  *peak_inflight_bytes = S->peak_inflight;
  *groups = S->groups;
  return CAC_OK;
}

int cac_server_submit(cac_server* S, const cac_query* q, uint64_t* id_out) {
  if (!S || !q) return CAC_EINVAL;
  if (q->kind != CAC_QUERY_CC && q->kind != CAC_QUERY_MINCUT) return CAC_EINVAL;
  cac::Pending p;
  p.q = *q;
  p.d_edges = nullptr;
  std::lock_guard<std::mutex> lk(S->mu);
  int v = CAC_OK;
  if (q->graph) {  // a registered graph: its size and device edges come from the registration
    auto it = S->graphs.find(q->graph);
    if (it == S->graphs.end() || it->second.dropped) {
      v = CAC_EINVAL;
    } else {
      p.q.n = it->second.n;
      p.q.m = it->second.m;
      p.q.edges = nullptr;
      p.d_edges = it->second.d_edges;
    }
  }
  if (v == CAC_OK) v = cac::validate(p.q, p.d_edges != nullptr);
  p.upload = p.scratch = 0;
  p.work = 0;
  if (v == CAC_OK) cac::footprint(p.q, p.d_edges != nullptr, &p.upload, &p.scratch);
  if (v == CAC_OK && p.q.kind == CAC_QUERY_MINCUT) {
    uint32_t t, T;
    cac::mincut_plan(p.q.n, p.q.m, p.q.p, p.q.trials, &t, &T);
    p.work = (uint64_t)T * std::max<uint64_t>(p.q.m, 1);
  }
  p.bytes = p.upload + p.scratch;
  p.id = S->next_id++;
  p.submit_s = S->now();
  if (id_out) *id_out = p.id;
  if (v != CAC_OK || p.bytes + S->resident > S->budget) {  // rejected now: reported through poll
    cac_query_result r = cac::blank(p);
    r.status = v != CAC_OK ? v : CAC_ENOMEM;
    r.admit_s = r.done_s = p.submit_s;
    S->done.push_back(r);
    S->cv.notify_all();
    return CAC_OK;
  }
  if (p.d_edges) ++S->graphs[q->graph].users;
  S->pending_work += p.work;
  S->pending.push_back(p);
  S->cv.notify_all();
  return CAC_OK;
}

int cac_server_register_graph(cac_server* S, const cac_edge* edges, uint64_t m, uint32_t n, uint64_t* graph_out) {
  if (!S || !graph_out || n < 1 || (m && !edges)) return CAC_EINVAL;
  const size_t bytes = cac::carve_size(std::max<uint64_t>(m, 1) * sizeof(cac_edge)) + 256;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->resident + bytes > S->budget) return CAC_ENOMEM;
  }
  if (hipSetDevice(S->device) != hipSuccess) return CAC_EDEVICE;
  void* d = nullptr;
  if (hipMalloc(&d, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return CAC_ENOMEM;
  }
  // ids, weights and the weight sum are checked once, here (as a min-cut
  // group checks an uploaded query)
  unsigned long long* chk = (unsigned long long*)((char*)d + bytes - 256);
  unsigned long long h[3] = {0, 0, 0};
  hipError_t e = m ? hipMemcpy(d, edges, m * sizeof(cac_edge), hipMemcpyHostToDevice) : hipSuccess;
  if (e == hipSuccess) e = hipMemcpy(chk, h, sizeof h, hipMemcpyHostToDevice);
  if (e == hipSuccess && m) {
    hipLaunchKernelGGL(cac::k_serve_check_w, dim3(cac::grid_for(m, 256, 2048)), dim3(256), 0, 0, (cac_edge*)d, m, n,
                       chk);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(h, chk, sizeof h, hipMemcpyDeviceToHost);
  int rc = e != hipSuccess ? CAC_EDEVICE
           : (h[0] & 1)   ? CAC_EINVAL
           : ((h[0] & 2) || h[1] >= (1ull << 61) || h[2] >= (1ull << 29)) ? CAC_ERANGE
                                                                          : CAC_OK;
  if (rc != CAC_OK) {
    (void)hipFree(d);
    return rc;
  }
  std::lock_guard<std::mutex> lk(S->mu);
  const uint64_t g = S->next_graph++;
  S->graphs[g] = cac::Resident{(cac_edge*)d, m, n, bytes, 0u, false};
  S->resident += bytes;
  *graph_out = g;
  return CAC_OK;
}

int cac_server_unregister_graph(cac_server* S, uint64_t graph) {
  if (!S) return CAC_EINVAL;
  cac_edge* d = nullptr;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->graphs.find(graph);
    if (it == S->graphs.end() || it->second.dropped) return CAC_EINVAL;
    if (it->second.users > 0) {  // freed by the worker that finishes its last query
      it->second.dropped = true;
      return CAC_OK;
    }
    d = it->second.d_edges;
    S->resident -= it->second.bytes;
    S->graphs.erase(it);
  }
  if (hipSetDevice(S->device) != hipSuccess) return CAC_EDEVICE;
  return hipFree(d) == hipSuccess ? CAC_OK : CAC_EDEVICE;
}

int cac_server_poll(cac_server* S, cac_query_result* out, uint32_t cap, uint32_t* count_out, int wait) {
  if (!S || !count_out || (cap && !out)) return CAC_EINVAL;
  std::unique_lock<std::mutex> lk(S->mu);
  if (wait)
    S->cv.wait(lk, [&] { return !S->done.empty() || (S->pending.empty() && S->inflight == 0); });
  const uint32_t k = (uint32_t)std::min<size_t>(cap, S->done.size());
  for (uint32_t i = 0; i < k; ++i) out[i] = S->done[i];
  S->done.erase(S->done.begin(), S->done.begin() + k);
  *count_out = k;
  return CAC_OK;
}

int cac_server_destroy(cac_server* S) {
  if (!S) return CAC_EINVAL;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv.notify_all();
  for (auto& t : S->workers) t.join();
  for (auto c : S->ctxs) cac_ctx_destroy(c);
  if (!S->graphs.empty() && hipSetDevice(S->device) == hipSuccess)
    for (auto& g : S->graphs) (void)hipFree(g.second.d_edges);
  delete S;
  return CAC_OK;
}

int cac_server_replay(cac_server* S, const cac_query* q, const double* arrival_s, uint32_t count,
                      cac_query_result* out) {
  if (!S || (count && (!q || !arrival_s || !out))) return CAC_EINVAL;
  const double start = S->now();
  std::vector<uint64_t> ids(count);
  std::vector<char> got(count, 0);
  uint32_t received = 0;
  uint64_t id0 = 0;
  std::vector<cac_query_result> buf(256);
  auto drain = [&](int wait) {
    uint32_t k = 0;
    cac_server_poll(S, buf.data(), (uint32_t)buf.size(), &k, wait);
    for (uint32_t i = 0; i < k; ++i) {
      const uint64_t idx = buf[i].id - id0;
      if (idx < count && !got[idx]) {
        out[idx] = buf[i];
        got[idx] = 1;
        ++received;
      }
    }
  };
  for (uint32_t i = 0; i < count; ++i) {
    for (;;) {  // wait for the arrival time, collecting finished results meanwhile
      const double dt = start + arrival_s[i] - S->now();
      if (dt <= 0) break;
      std::this_thread::sleep_for(std::chrono::duration<double>(std::min(dt, 0.0005)));
      drain(0);
    }
    uint64_t id = 0;
    int rc = cac_server_submit(S, &q[i], &id);
    if (rc != CAC_OK) return rc;
    if (i == 0) id0 = id;
    ids[i] = id;
    if (id != id0 + i) return CAC_EINVAL;  // another thread submitted meanwhile
  }
  while (received < count) drain(1);
  return CAC_OK;
}

}  // extern "C"
