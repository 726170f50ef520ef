// cac_common.h -- shared device helpers for the gfx950 contraction-core kernels.
//
// Counter-based PRNG, uniform-integer rule and launch helpers.  The PRNG replaces
// the reference's sequential sitmo::prng_engine stream (lib/prng_engine.hpp:195-212):
// every draw is addressed by (seed, domain, counter words), so trials, recursion
// nodes and sample draws can run in any order on the device and still match the
// CPU restatement bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cac.h"

namespace cac {

// Domain tags occupy key word 1 so the streams of different consumers never overlap.
constexpr uint32_t kDomRmat = 0x524D4154u;    // 'RMAT'  synthetic R-MAT edges
constexpr uint32_t kDomSample = 0x53414D50u;  // 'SAMP'  weighted edge sampling draws
constexpr uint32_t kDomKs = 0x4B415247u;      // 'KARG'  dense Karger-Stein picks
constexpr uint32_t kDomApprox = 0x41505258u;  // 'APRX'  approximate-cut subsampling
constexpr uint32_t kDomGen = 0x47454E52u;     // 'GENR'  other synthetic generators

struct u32x4 {
  uint32_t x, y, z, w;
};

// Philox-4x32-10 (Salmon et al., SC'11): 10 rounds, Weyl key schedule.
__host__ __device__ inline u32x4 philox4x32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                            uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c0 = n0;
    c1 = (uint32_t)p1;
    c2 = n2;
    c3 = (uint32_t)p0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return {c0, c1, c2, c3};
}

__host__ __device__ inline uint64_t lo64(const u32x4& r) { return ((uint64_t)r.y << 32) | r.x; }
__host__ __device__ inline uint64_t hi64(const u32x4& r) { return ((uint64_t)r.w << 32) | r.z; }

// Uniform integer in [1, W]: 1 + floor(x * W / 2^64).  Replaces
// std::uniform_int_distribution<u64>(1, W) (src/SequentialSquareRootCut.cpp:77-81).
__host__ __device__ inline uint64_t uniform1(uint64_t x, uint64_t W) {
#if defined(__HIP_DEVICE_COMPILE__)
  return 1u + __umul64hi(x, W);
#else
  return 1u + (uint64_t)(((unsigned __int128)x * W) >> 64);
#endif
}

// Barrier after which global-memory writes of this workgroup are visible to all
// of its waves.  __syncthreads() alone only orders LDS on gfx950: the compiler
// does not wait for outstanding vector stores before s_barrier, and the CU's
// vector L1 may still hold lines loaded before another wave's store.  The
// agent-scope acquire fence waits for this wave's memory operations
// (vmcnt(0)) and invalidates the L1, so loads after the barrier read L2.
__device__ inline void block_sync_global() {
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  __syncthreads();
}

inline unsigned grid_for(uint64_t work, unsigned block, unsigned cap) {
  uint64_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

}  // namespace cac
