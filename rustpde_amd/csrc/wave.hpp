// Cross-lane sums inside one wavefront through DPP moves and lane reads: no LDS round trip.
#pragma once
#include <hip/hip_runtime.h>

namespace rbc {
namespace wave {

// DPP controls (GFX9 encoding): row_shl:n moves lane l + n of the lane's row of 16 into
// lane l (lanes whose source leaves the row keep the 0 given as `old`); wave_shl:1 moves
// lane l + 1 of the wavefront (lane 63 keeps the 0)
constexpr int ROW_SHL = 0x100, WAVE_SHL1 = 0x130;

template <int CTRL, typename V>
__device__ __forceinline__ V dpp(V x) {
  if constexpr (sizeof(V) == 4) {
    return __builtin_bit_cast(V, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xf, 0xf, false));
  } else {
    const long long u = __builtin_bit_cast(long long, x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)u, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(V, ((long long)hi << 32) | (unsigned)lo);
  }
}

template <typename V>
__device__ __forceinline__ V lane_value(V x, int l) {
  if constexpr (sizeof(V) == 4) {
    return __builtin_bit_cast(V, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
  } else {
    const long long u = __builtin_bit_cast(long long, x);
    const int lo = __builtin_amdgcn_readlane((int)u, l), hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
    return __builtin_bit_cast(V, ((long long)hi << 32) | (unsigned)lo);
  }
}

// inclusive suffix sums over the 64 lanes, s_l = x_l + x_{l+1} + ... + x_63: suffix sums
// inside each row of 16 (steps 1, 2, 4, 8), then the totals of the rows above (read from
// their first lanes) added per row; excl receives s_{l+1} (0 for lane 63)
template <typename V>
__device__ __forceinline__ V suffix_sum(V x, int lane, V& excl) {
  x = x + dpp<ROW_SHL + 1>(x);
  x = x + dpp<ROW_SHL + 2>(x);
  x = x + dpp<ROW_SHL + 4>(x);
  x = x + dpp<ROW_SHL + 8>(x);
  const V t1 = lane_value(x, 16), t2 = lane_value(x, 32), t3 = lane_value(x, 48);
  const V s2 = t2 + t3, s1 = t1 + s2;
  const int row = lane >> 4;
  const V above = row == 0 ? s1 : row == 1 ? s2 : row == 2 ? t3 : (V)0;
  x = x + above;
  excl = dpp<WAVE_SHL1>(x);
  return x;
}

}  // namespace wave
}  // namespace rbc
