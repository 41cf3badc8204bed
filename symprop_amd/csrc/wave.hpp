// Wavefront reductions for the device code (HIP sources only).
#pragma once

#include <hip/hip_runtime.h>

namespace spamd {

// One DPP move of a double (two 32-bit halves); C is the DPP control word.
template <int C>
__device__ inline double dpp_move(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, C, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), C, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ inline double read_lane(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over the 64 lanes of a full wave, the same value in every lane, in a
// fixed order: within each quad (quad_perm [1,0,3,2], then [2,3,0,1]), within
// each row of 16 (row_ror 4, then 8), then the four rows' sums.  DPP moves
// are VALU operations; shuffles through the LDS permute unit would cost an
// LDS round trip per step.
__device__ inline double wave_sum(double x) {
  x += dpp_move<0xb1>(x);
  x += dpp_move<0x4e>(x);
  x += dpp_move<0x124>(x);
  x += dpp_move<0x128>(x);
  return (read_lane(x, 0) + read_lane(x, 16)) + (read_lane(x, 32) + read_lane(x, 48));
}

}  // namespace spamd
