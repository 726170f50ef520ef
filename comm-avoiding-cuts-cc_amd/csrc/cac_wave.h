// cac_wave.h -- wavefront (64-lane) reductions and scans by DPP row moves.
//
// Every function here is called by whole wavefronts (no lane of the wave
// inactive at the call).  Cross-lane steps are DPP moves inside rows of 16
// lanes and the row broadcasts of lanes 15 and 31: VALU data moves of a few
// cycles each, where LDS permutes (what __shfl compiles to) take a round trip
// through the LDS crossbar per step.  The kernels that are bound by a chain of
// such steps (the dense contraction's picks, a sampling workgroup's rounds) run
// a dozen of them per iteration.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cac {

// DPP controls (gfx9 encoding).
constexpr int kDppQuad1032 = 0xb1;  // quad_perm [1,0,3,2]
constexpr int kDppQuad2301 = 0x4e;  // quad_perm [2,3,0,1]
constexpr int kDppRowShr1 = 0x111;
constexpr int kDppRowShr2 = 0x112;
constexpr int kDppRowShr4 = 0x114;
constexpr int kDppRowShr8 = 0x118;
constexpr int kDppRowRor4 = 0x124;
constexpr int kDppRowRor8 = 0x128;
constexpr int kDppBcast15 = 0x142;  // lane 15 of each row into the next row
constexpr int kDppBcast31 = 0x143;  // lane 31 into rows 2 and 3

// One DPP move; lanes whose source lane lies outside its row, or whose own row
// is masked off, get `old`.
template <int Ctrl, int RowMask>
__device__ inline uint32_t dpp_u32(uint32_t v, uint32_t old = 0) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, Ctrl, RowMask, 0xf, false);
}
template <int Ctrl, int RowMask>
__device__ inline uint64_t dpp_u64(uint64_t v, uint64_t old = 0) {
  const uint32_t lo = dpp_u32<Ctrl, RowMask>((uint32_t)v, (uint32_t)old);
  const uint32_t hi = dpp_u32<Ctrl, RowMask>((uint32_t)(v >> 32), (uint32_t)(old >> 32));
  return ((uint64_t)hi << 32) | lo;
}
template <int Ctrl, int RowMask>
__device__ inline uint32_t dpp_move(uint32_t v, uint32_t old) {
  return dpp_u32<Ctrl, RowMask>(v, old);
}
template <int Ctrl, int RowMask>
__device__ inline uint64_t dpp_move(uint64_t v, uint64_t old) {
  return dpp_u64<Ctrl, RowMask>(v, old);
}

__device__ inline uint32_t readlane_u32(uint32_t v, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, l); }
__device__ inline uint64_t readlane_u64(uint64_t v, int l) {
  const uint32_t lo = readlane_u32((uint32_t)v, l);
  const uint32_t hi = readlane_u32((uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}
__device__ inline uint32_t readlane(uint32_t v, int l) { return readlane_u32(v, l); }
__device__ inline uint64_t readlane(uint64_t v, int l) { return readlane_u64(v, l); }

// Inclusive prefix sum over the wave (T = uint32_t or uint64_t, wrapping):
// Hillis-Steele inside each row (row_shr 1, 2, 4, 8; a lane with no source adds
// 0), then the broadcasts of lane 15 into rows 1 and 3 and of lane 31 into
// rows 2 and 3.
template <class T>
__device__ inline T wave_incl_scan(T x) {
  x += dpp_move<kDppRowShr1, 0xf>(x, (T)0);
  x += dpp_move<kDppRowShr2, 0xf>(x, (T)0);
  x += dpp_move<kDppRowShr4, 0xf>(x, (T)0);
  x += dpp_move<kDppRowShr8, 0xf>(x, (T)0);
  x += dpp_move<kDppBcast15, 0xa>(x, (T)0);
  x += dpp_move<kDppBcast31, 0xc>(x, (T)0);
  return x;
}
// Wave total, returned to every lane.
template <class T>
__device__ inline T wave_sum(T x) {
  return readlane(wave_incl_scan(x), 63);
}

// Wave minimum / maximum (unsigned), returned to every lane: inside each row by
// two quad permutations and two row rotations, then across rows by the two
// broadcasts, which leave the result in lane 63; lanes that receive nothing
// keep the operation's identity.
template <class T>
__device__ inline T wave_min(T x) {
  const T id = (T)~(T)0;
  T o = dpp_move<kDppQuad1032, 0xf>(x, id);
  x = o < x ? o : x;
  o = dpp_move<kDppQuad2301, 0xf>(x, id);
  x = o < x ? o : x;
  o = dpp_move<kDppRowRor4, 0xf>(x, id);
  x = o < x ? o : x;
  o = dpp_move<kDppRowRor8, 0xf>(x, id);
  x = o < x ? o : x;
  o = dpp_move<kDppBcast15, 0xa>(x, id);
  x = o < x ? o : x;
  o = dpp_move<kDppBcast31, 0xc>(x, id);
  x = o < x ? o : x;
  return readlane(x, 63);
}
template <class T>
__device__ inline T wave_max(T x) {
  const T id = (T)0;
  T o = dpp_move<kDppQuad1032, 0xf>(x, id);
  x = o > x ? o : x;
  o = dpp_move<kDppQuad2301, 0xf>(x, id);
  x = o > x ? o : x;
  o = dpp_move<kDppRowRor4, 0xf>(x, id);
  x = o > x ? o : x;
  o = dpp_move<kDppRowRor8, 0xf>(x, id);
  x = o > x ? o : x;
  o = dpp_move<kDppBcast15, 0xa>(x, id);
  x = o > x ? o : x;
  o = dpp_move<kDppBcast31, 0xc>(x, id);
  x = o > x ? o : x;
  return readlane(x, 63);
}

}  // namespace cac
