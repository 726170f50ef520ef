// cac_sort.h -- device radix sort of (u64 key, u64 value) pairs and the
// segmented sum of equal keys (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "cac_internal.h"

namespace cac {

// Stable LSD radix sort of n pairs by the low `bits` key bits (8 bits per
// pass).  keys/vals hold the input; kalt/valt are buffers of n pairs each.
// On return the sorted pairs are in (keys, vals) when *in_alt is false and in
// (kalt, valt) when it is true.  temp holds radix_sort_temp_bytes(n) bytes.
size_t radix_sort_temp_bytes(uint64_t n);
int radix_sort_pairs(cac_ctx* ctx, uint64_t* keys, uint64_t* vals, uint64_t* kalt, uint64_t* valt, uint64_t n,
                     int bits, void* temp, bool* in_alt, hipStream_t s);

// Segmented sum of a sorted (key, value) array: out_keys[j * stride] and
// out_vals[j * stride] = the j-th distinct key that is kept and the sum of its
// values; a key is dropped when (key & drop_mask) == drop_mask.  head_idx[i]
// (n + 1 entries, u32: n < 2^32) receives the exclusive count of kept distinct
// keys before position i, so the output range of a sorted key range [p, q) is
// [head_idx[p], head_idx[q]).  temp holds segmented_sum_temp_bytes(n) bytes.
// Integer sums: the result does not depend on the order of the additions.
size_t segmented_sum_temp_bytes(uint64_t n);
int segmented_sum(cac_ctx* ctx, const uint64_t* keys, const uint64_t* vals, uint64_t n, uint64_t drop_mask,
                  uint64_t* out_keys, uint64_t* out_vals, uint32_t stride, uint32_t* head_idx, void* temp,
                  hipStream_t s);

}  // namespace cac
