// cac_trial.h -- one sqrt-cut trial's device scratch, shared by the
// workgroup-per-trial sampling kernels and the wide (grid-per-trial) ones
// (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "cac_internal.h"

namespace cac {

// One sampling job = one sqrt-cut trial of one graph; a launch runs one
// workgroup per job, so trials of different graphs share a launch.
struct TrialScratch {
  const cac_edge* edges;  // the job's graph (ids local to it)
  uint64_t m;
  uint32_t n, t;          // vertices, intermediate size
  uint32_t trial, slot;   // Philox trial id, result slot
  uint32_t seed;          // Philox key word 0 of the trial
  uint32_t dpr;           // draws per round: 0 = |V| (low-concurrency rule), else that many
  uint32_t* lab;          // [n]
  uint32_t* par;          // [n]
  uint64_t* pre;          // [m]
  uint2* draws;           // [n]
  int64_t* mat;           // [t x t]
  int64_t* rsum;          // [t] row sums of mat (may be null)
  uint32_t* nv_out;       // [1]
  // [1] of the job's graph, nonzero when an edge is a loop or two weights
  // differ (may be null: then every round builds the prefix)
  const uint32_t* mixed;
};

// Device bytes of the per-trial state of the wide form beside TrialScratch's
// buffers (tile totals and the round's control words).
size_t wide_trial_bytes(uint64_t m, uint32_t n);

// Wide trials: every trial of h (device copies at d_scr, k = h.size()) runs its
// sampling rounds as grid-wide kernels over its edge list instead of in one
// workgroup, with the same draws, draw-prefix lengths and canonical labels.
// work holds k * wide_trial_bytes(m, n) bytes carved in order (the caller's
// sizes).  On return (queued on s, after two host synchronisations per round)
// lab holds the canonical labels and *nv_out the super-vertex count.
int wide_sample(cac_ctx* ctx, const std::vector<TrialScratch>& h, const TrialScratch* d_scr, void* work,
                hipStream_t s);
// Then the t x t matrix, row sums and bound of each trial, as k_densify_trials
// writes them: bounds[i] (i indexes h) and results[h[i].slot].
int wide_densify(cac_ctx* ctx, const std::vector<TrialScratch>& h, const TrialScratch* d_scr, void* work,
                 int64_t* bounds, int64_t* results, hipStream_t s);

// Smallest edge count that runs a trial in the wide form (the context's
// setting or the built-in 2^20).
uint64_t wide_min_edges(const cac_ctx* ctx);

}  // namespace cac
