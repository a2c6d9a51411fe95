// Wavefront reductions on DPP lane moves (CDNA: 64-lane waves, 16-lane rows).
//
// quad_perm [1,0,3,2] and [2,3,0,1] combine lanes within a quad, row_half_mirror
// (lane i <- 7 - i inside each 8-lane half-row) joins two quads, row_mirror
// (lane i <- 15 - i inside each row) joins two half-rows; four readlanes join
// the rows.  Each step combines the same two operands on both partner lanes, so
// every lane of the group ends with a bitwise identical result (commutative
// operations), which keeps data-dependent control flow uniform.
#pragma once

#include <hip/hip_runtime.h>

namespace fpls {

namespace dpp {
constexpr int QUAD_XOR1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int QUAD_XOR2 = 0x4E;  // quad_perm [2,3,0,1]
constexpr int HALF_MIRROR = 0x141;
constexpr int ROW_MIRROR = 0x140;
}  // namespace dpp

template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ int dpp_mov(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

// sums / products over aligned groups of 8 lanes
__device__ __forceinline__ double grp8_sum(double v) {
  v += dpp_mov<dpp::QUAD_XOR1>(v);
  v += dpp_mov<dpp::QUAD_XOR2>(v);
  v += dpp_mov<dpp::HALF_MIRROR>(v);
  return v;
}
__device__ __forceinline__ double grp8_prod(double v) {
  v *= dpp_mov<dpp::QUAD_XOR1>(v);
  v *= dpp_mov<dpp::QUAD_XOR2>(v);
  v *= dpp_mov<dpp::HALF_MIRROR>(v);
  return v;
}
__device__ __forceinline__ int grp8_isum(int v) {
  v += dpp_mov<dpp::QUAD_XOR1>(v);
  v += dpp_mov<dpp::QUAD_XOR2>(v);
  v += dpp_mov<dpp::HALF_MIRROR>(v);
  return v;
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// whole-wave sum / max; the result is wave-uniform
__device__ __forceinline__ double wave_sum_d(double v) {
  v = grp8_sum(v);
  v += dpp_mov<dpp::ROW_MIRROR>(v);
  return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}
__device__ __forceinline__ double wave_max_d(double v) {
  v = fmax(v, dpp_mov<dpp::QUAD_XOR1>(v));
  v = fmax(v, dpp_mov<dpp::QUAD_XOR2>(v));
  v = fmax(v, dpp_mov<dpp::HALF_MIRROR>(v));
  v = fmax(v, dpp_mov<dpp::ROW_MIRROR>(v));
  return fmax(fmax(readlane_d(v, 0), readlane_d(v, 16)), fmax(readlane_d(v, 32), readlane_d(v, 48)));
}

}  // namespace fpls
