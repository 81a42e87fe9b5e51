// ec_wave.hpp -- wavefront sums of doubles with DPP lane moves (VALU only,
// no LDS traffic), in a fixed tree: two quad swaps, the half-row and row
// mirrors, then the broadcasts of lanes 15 and 31 into the following rows.
// The total ends in lane 63.  Every lane of the wavefront must be active.
#pragma once

#include <hip/hip_runtime.h>

namespace ec {

// The value of lane sel(i) for each lane i under DPP control CTRL; rows
// outside ROW_MASK read 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v)
{
  const long long b = __double_as_longlong(v);
  const int lo = (int)(unsigned)(b & 0xffffffffLL);
  const int hi = (int)(b >> 32);
  const int lo2 = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROW_MASK, 0xf, false);
  const int hi2 = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double(((long long)hi2 << 32) | (long long)(unsigned)lo2);
}

// The same move for a double whose rows outside ROW_MASK read `id`, and for
// a 32-bit integer.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64_id(double v, double id)
{
  const long long b = __double_as_longlong(v), o = __double_as_longlong(id);
  const int lo2 = __builtin_amdgcn_update_dpp((int)(unsigned)(o & 0xffffffffLL),
                                              (int)(unsigned)(b & 0xffffffffLL), CTRL, ROW_MASK,
                                              0xf, false);
  const int hi2 = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), CTRL, ROW_MASK, 0xf,
                                              false);
  return __longlong_as_double(((long long)hi2 << 32) | (long long)(unsigned)lo2);
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_i32(int v)
{
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xf, false);
}

// Moves under a control that gives every lane a lane of its own row (quad
// permutes, row mirrors) with every lane active: no lane keeps an old value,
// so none is supplied and no copy is made.
template <int CTRL>
__device__ __forceinline__ double dpp_f64_all(double v)
{
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)(b & 0xffffffffLL), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int LANE>
__device__ __forceinline__ double readlane_f64(double v)
{
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffLL), LANE);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), LANE);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int CTRL>
__device__ __forceinline__ int dpp_i32_all(int v)
{
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, false);
}

__device__ __forceinline__ double wave_sum_lane63(double v)
{
  v += dpp_f64<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]: pairs
  v += dpp_f64<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]: quads
  v += dpp_f64<0x141, 0xF>(v);  // row_half_mirror: 8-lane halves
  v += dpp_f64<0x140, 0xF>(v);  // row_mirror: 16-lane rows
  v += dpp_f64<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_f64<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}

// Swizzle of a double within 32-lane halves (bit-mask mode, xor XM).
template <int XM>
__device__ __forceinline__ double swizzle_xor_f64(double v)
{
  constexpr int pattern = 0x1F | (XM << 10);  // and 0x1F, or 0, xor XM
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_swizzle((int)b, pattern);
  const int hi = __builtin_amdgcn_ds_swizzle((int)(b >> 32), pattern);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sums each of v[0..7] over the wavefront with a reduce-scatter butterfly:
// three exchanges halve the values each lane holds (partners across 32
// lanes, 16 lanes and the mirrored half of each 16-lane row), three DPP
// steps finish the sum over each 8-lane group.  Returns, in every lane, the
// total of value (lane >> 3) & 7.  The summation order is fixed.
__device__ __forceinline__ double wave_sum8(const double (&v)[8])
{
  const int lane = threadIdx.x & 63;
  double a[4], b[2];
  const bool h32 = (lane & 32) != 0, h16 = (lane & 16) != 0, h8 = (lane & 8) != 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double keep = h32 ? v[4 + i] : v[i];
    const double send = h32 ? v[i] : v[4 + i];
    a[i] = keep + __shfl_xor(send, 32);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double keep = h16 ? a[2 + i] : a[i];
    const double send = h16 ? a[i] : a[2 + i];
    b[i] = keep + swizzle_xor_f64<16>(send);
  }
  double c = (h8 ? b[1] : b[0]) + dpp_f64<0x140, 0xF>(h8 ? b[0] : b[1]);  // row_mirror
  c += dpp_f64<0xB1, 0xF>(c);   // quad_perm [1,0,3,2]
  c += dpp_f64<0x4E, 0xF>(c);   // quad_perm [2,3,0,1]
  c += dpp_f64<0x141, 0xF>(c);  // row_half_mirror
  return c;
}
}  // namespace ec
