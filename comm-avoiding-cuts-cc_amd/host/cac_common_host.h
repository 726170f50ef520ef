// cac_common_host.h -- the PRNG spec for host-only code (same words as the device).
#pragma once

#include <cstdint>

namespace cac {

constexpr uint32_t kDomGen = 0x47454E52u;  // 'GENR'

struct u32x4 {
  uint32_t x, y, z, w;
};

// Philox-4x32-10 (Salmon et al., SC'11).
inline u32x4 philox4x32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
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

}  // namespace cac
