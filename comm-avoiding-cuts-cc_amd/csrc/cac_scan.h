// cac_scan.h -- device-wide exclusive prefix sum (reduce / scan-partials /
// rescan), templated on value type, input loader and output writer.
//
// Used for the dense ranks of canonical CC labels and for the integer weight
// prefix that drives edge sampling.  DPP wave scans inside a block, LDS for the
// per-wave totals, one partial per 4096-element tile.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cac_wave.h"

namespace cac {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

template <class T>
__device__ inline T wave_inclusive_scan(T x) {
  return wave_incl_scan<T>(x);
}

// Exclusive scan of one value per thread across the block; returns the block total.
template <class T>
__device__ inline T block_exclusive_scan(T x, T* lds_wave /* [kScanBlock/64] */, T& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T inc = wave_inclusive_scan(x);
  if (lane == 63) lds_wave[wave] = inc;
  __syncthreads();
  T off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kScanBlock / 64; ++w) {
    T v = lds_wave[w];
    if (w < wave) off += v;
    tot += v;
  }
  __syncthreads();
  total = tot;
  return off + inc - x;
}

template <class T, class Load>
__global__ void __launch_bounds__(kScanBlock) k_scan_reduce(Load load, uint64_t n, T* partial) {
  __shared__ T lds[kScanBlock / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  T s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    uint64_t i = base + (uint64_t)k * kScanBlock + threadIdx.x;
    if (i < n) s += load(i);
  }
  T tot;
  (void)block_exclusive_scan(s, lds, tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Single block: exclusive scan of the tile partials in place; total to *sum.
template <class T>
__global__ void __launch_bounds__(kScanBlock) k_scan_partials(T* partial, uint64_t np, T* sum) {
  __shared__ T lds[kScanBlock / 64];
  T carry = 0;
  for (uint64_t base = 0; base < np; base += kScanBlock) {
    uint64_t i = base + threadIdx.x;
    T x = i < np ? partial[i] : T(0);
    T tot;
    T ex = block_exclusive_scan(x, lds, tot);
    if (i < np) partial[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && sum) *sum = carry;
}

// Rescan: each thread owns kScanItems consecutive elements of the tile.
template <class T, class Load, class Store>
__global__ void __launch_bounds__(kScanBlock) k_scan_apply(Load load, Store store, uint64_t n,
                                                          const T* partial) {
  __shared__ T lds[kScanBlock / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  T v[kScanItems];
  T s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    uint64_t i = base + k;
    v[k] = i < n ? load(i) : T(0);
    s += v[k];
  }
  T tot;
  T ex = block_exclusive_scan(s, lds, tot) + partial[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    uint64_t i = base + k;
    if (i < n) store(i, ex, v[k]);
    ex += v[k];
  }
}

inline uint64_t scan_tiles(uint64_t n) { return (n + kScanTile - 1) / kScanTile; }

// Exclusive scan: store(i, prefix_before_i, value_i).  partial needs scan_tiles(n)
// entries; *sum (device pointer, may be null) receives the total.
template <class T, class Load, class Store>
inline hipError_t device_exclusive_scan(Load load, Store store, uint64_t n, T* partial, T* sum,
                                        hipStream_t s) {
  uint64_t nt = scan_tiles(n);
  if (nt == 0) {  // empty input: only the (zero) total is written
    hipLaunchKernelGGL((k_scan_partials<T>), dim3(1), dim3(kScanBlock), 0, s, partial, (uint64_t)0, sum);
    return hipGetLastError();
  }
  hipLaunchKernelGGL((k_scan_reduce<T, Load>), dim3((unsigned)nt), dim3(kScanBlock), 0, s, load, n, partial);
  hipLaunchKernelGGL((k_scan_partials<T>), dim3(1), dim3(kScanBlock), 0, s, partial, nt, sum);
  hipLaunchKernelGGL((k_scan_apply<T, Load, Store>), dim3((unsigned)nt), dim3(kScanBlock), 0, s, load, store, n,
                     (const T*)partial);
  return hipGetLastError();
}

}  // namespace cac
