// Wavefront helpers of the element kernels (device code).
#pragma once

#include <hip/hip_runtime.h>

namespace pxg
{

// a double moved across lanes by a DPP pattern within rows of 16 lanes
template <int CTRL>
__device__ inline double dppMove(double x)
{
  unsigned long long const b = (unsigned long long)__double_as_longlong(x);
  int const lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
  int const hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// sum over the 16 lanes of a row (pairs, quads, halves, rows: every lane
// gets the same bits, IEEE addition being commutative)
__device__ inline double rowSum16(double x)
{
  x += dppMove<0xB1>(x);  // quad_perm [1, 0, 3, 2]
  x += dppMove<0x4E>(x);  // quad_perm [2, 3, 0, 1]
  x += dppMove<0x141>(x); // row_half_mirror
  x += dppMove<0x140>(x); // row_mirror
  return x;
}

} // namespace pxg
