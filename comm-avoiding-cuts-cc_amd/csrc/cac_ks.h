// cac_ks.h -- host-side interface of the dense Karger-Stein engine (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "cac_internal.h"

namespace cac {

// One trial's root: a dense n x n device matrix (row stride n) with its row
// sums, and where its raw cut goes.  Several roots may share one matrix
// (independent trials).  Its draws are keyed by (seed, trial, node id).
struct KsRoot {
  const int64_t* mat;
  const int64_t* rsum;
  uint32_t n, trial, slot;
  int64_t bound;     // the trial's upper bound (minimum weighted degree)
  uint32_t seed;     // Philox key word 0 of the trial
  uint32_t nid = 1;  // recursion node id of the matrix (1: the trial's root)
  bool w32 = false;  // total weight below 2^32: every node of the recursion fits 32-bit sums
};

// Row sums of a dense device matrix into d_rsum and their minimum into *d_min.
int ks_row_sums(cac_ctx* ctx, const int64_t* d_adj, uint32_t n, int64_t* d_rsum, int64_t* d_min, hipStream_t s);
// Upper estimate of the device bytes one root holds at once in the two level
// buffers (even depths in one, odd depths in the other, each grown to its
// largest level: children arrays and row pools).
size_t ks_live_bytes_per_root(uint32_t n);
// Runs one trial per root (roots of different seeds share launches);
// d_results[slot] (initialised to the root's bound) receives min(bound, leaves).
int ks_run(cac_ctx* ctx, const std::vector<KsRoot>& roots, int64_t* d_results, hipStream_t s);

// Host forms of the reference's size and trial-count rules (same floating point).
uint32_t ks_target_host(uint32_t s);
uint32_t ks_trials_host(uint32_t n, double p);
uint32_t sqrt_intermediate_size(uint32_t n, uint64_t m);
uint32_t sqrt_trials_host(uint32_t n, uint64_t m, double p);
// One sqrt-cut job: trial 'trial' of the graph (edges, m, n) shrunk to t
// super-vertices with the draws of Philox key `seed`, raw cut to
// d_results[slot].  Jobs of different graphs and seeds share launches.
struct SqrtJob {
  const cac_edge* edges;
  uint64_t m;
  uint32_t n, t, trial, slot, seed;
};
constexpr size_t kMaxJobsPerChunk = 8192;
// Jobs of any mix of graphs, in chunks of shared launches whose estimated
// device bytes stay within budget (0: 40% of the free device memory).
// by_size: the workgroup-form jobs are chunked largest graph first (see the
// definition for when), not in the order given.
int sqrt_cut_batch(cac_ctx* ctx, const std::vector<SqrtJob>& jobs, int64_t* d_results, hipStream_t s,
                   size_t budget = 0, bool by_size = false);
// Estimated device bytes of one job: sampling scratch (dpr: draws per round,
// 0 = |V|), Karger-Stein levels.
size_t job_sample_bytes(const SqrtJob& j, uint32_t dpr = 0);
// Draws per round of the high-concurrency sampling rule for n0 vertices.
uint32_t hc_draws_host(uint32_t n0);
size_t job_ks_bytes(const SqrtJob& j);
// Sqrt-cut trials on a device edge list: raw cut of trial_ids[i] to d_results[i].
int sqrt_cut_run(cac_ctx* ctx, const cac_edge* d_edges, uint64_t m, uint32_t n, uint32_t t, uint32_t seed,
                 const std::vector<uint32_t>& trial_ids, int64_t* d_results, hipStream_t s);
// Sparse Karger-Stein trials on a checked device edge list (u < v, 1 <= w < 2^40):
// raw cut of trial_ids[i] to d_results[i], which must already hold the bound.
int ks_sparse_run(cac_ctx* ctx, const cac_edge* d_edges, uint64_t m, uint32_t n, uint32_t seed,
                  const std::vector<uint32_t>& trial_ids, int64_t bound, int64_t* d_results, hipStream_t s);
// Sparse trials of possibly different graphs side by side: root i is trial
// `trial` of (edges, m, n) from node 1; d_results[slot] must hold its bound.
struct SparseRoot {
  const cac_edge* edges;
  uint64_t m;
  uint32_t n, trial, slot;
  int64_t bound;
  uint32_t seed;
};
int ks_sparse_roots(cac_ctx* ctx, const std::vector<SparseRoot>& roots, int64_t* d_results, hipStream_t s);
int ks_check_matrix(cac_ctx* ctx, const int64_t* adj, uint32_t n);

// One sampling contraction (cac_sample_contract): canonical labels and the merged
// k x k matrix stay on the device until sample_out_free.
struct SampleOut {
  uint32_t* lab = nullptr;
  int64_t* mat = nullptr;
  uint32_t k = 0;
  void* arena = nullptr;
};
int sample_contract_one(cac_ctx* ctx, const cac_edge* d_edges, uint64_t m, uint32_t n, uint32_t t, uint32_t seed,
                        uint32_t trial, SampleOut* out, hipStream_t s);
void sample_out_free(SampleOut* o);

}  // namespace cac
