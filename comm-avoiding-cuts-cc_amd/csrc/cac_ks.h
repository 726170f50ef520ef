// cac_ks.h -- host-side interface of the dense Karger-Stein engine (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "cac_internal.h"

namespace cac {

// One trial's root: a dense n x n device matrix (row stride n) and where its raw
// cut goes.  Several roots may share one matrix (independent trials).
struct KsRoot {
  const int64_t* mat;
  uint32_t n, trial, slot;
};

// Minimum weighted degree of a dense device matrix into *d_bound (device).
int ks_min_degree(cac_ctx* ctx, const int64_t* d_adj, uint32_t n, int64_t* d_bound, hipStream_t s);
// Runs one trial per root; d_bounds[slot] holds each trial's bound and
// d_results[slot] (initialised to the bound) receives min(bound, leaves).
int ks_run(cac_ctx* ctx, const std::vector<KsRoot>& roots, uint32_t seed, const int64_t* d_bounds,
           int64_t* d_results, hipStream_t s);
int ks_check_matrix(cac_ctx* ctx, const int64_t* adj, uint32_t n);

}  // namespace cac
