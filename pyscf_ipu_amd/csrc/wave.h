// Cross-lane building blocks for 64-lane wavefronts on gfx950: DPP moves
// (VALU, no LDS traffic) and the permlane swaps.
#pragma once
#include <hip/hip_runtime.h>

namespace pscf {

// DPP move of v (control CTRL, e.g. row_shr:k = 0x110 + k, row_mirror 0x140,
// row_bcast:15 / :31 = 0x142 / 0x143, quad_perm codes below 0x100).  Lanes
// outside ROW_MASK, or whose source lane is outside the row, receive `old`
// (0 for the floating-point forms).
template <int CTRL, int ROW_MASK>
__device__ inline int dpp_i(int v, int old) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, ROW_MASK, 0xf, false);
}
template <int CTRL, int ROW_MASK>
__device__ inline float dpp_t(float v) {
  return __int_as_float(dpp_i<CTRL, ROW_MASK>(__float_as_int(v), 0));
}
template <int CTRL, int ROW_MASK>
__device__ inline double dpp_t(double v) {
  const int lo = dpp_i<CTRL, ROW_MASK>(__double2loint(v), 0);
  const int hi = dpp_i<CTRL, ROW_MASK>(__double2hiint(v), 0);
  return __hiloint2double(hi, lo);
}

// One halving level of a transpose-reduction: vectors x[i] and x[i + H] of
// lane l and its DPP partner (which differs from l in lane bit `bit`) become
// one vector, lanes with the bit clear keeping the sums of x[i], lanes with it
// set the sums of x[i + H].
template <int CTRL, int H, typename T>
__device__ inline void tr_level(T (&x)[16], bool bit) {
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const T keep = bit ? x[i + H] : x[i];
    const T send = bit ? x[i] : x[i + H];
    x[i] = keep + dpp_t<CTRL, 0xf>(send);
  }
}
__device__ inline float xor16_sum(float v) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_int(v), __float_as_int(v), false, false);
  return __int_as_float(r[0]) + __int_as_float(r[1]);
}
__device__ inline float xor32_sum(float v) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_int(v), __float_as_int(v), false, false);
  return __int_as_float(r[0]) + __int_as_float(r[1]);
}
__device__ inline double xor16_sum(double v) {
  const auto h = __builtin_amdgcn_permlane16_swap(__double2hiint(v), __double2hiint(v), false, false);
  const auto l = __builtin_amdgcn_permlane16_swap(__double2loint(v), __double2loint(v), false, false);
  return __hiloint2double(h[0], l[0]) + __hiloint2double(h[1], l[1]);
}
__device__ inline double xor32_sum(double v) {
  const auto h = __builtin_amdgcn_permlane32_swap(__double2hiint(v), __double2hiint(v), false, false);
  const auto l = __builtin_amdgcn_permlane32_swap(__double2loint(v), __double2loint(v), false, false);
  return __hiloint2double(h[0], l[0]) + __hiloint2double(h[1], l[1]);
}

// Sums each of 16 vectors over the 64 lanes at once; lane l returns the total
// of vector x[l & 15].  Pairs i <-> 15-i (row_mirror), i <-> 7-i
// (row_half_mirror), i <-> i^2 and i <-> i^1 (quad_perm) halve the vector
// count while keeping the partners' vector sets equal, then the four rows of
// 16 lanes are added with the permlane swaps: 49 VALU operations for 16
// reductions instead of about 160 for 16 separate ones.
template <typename T>
__device__ inline T transpose_reduce16(T (&x)[16], int lane) {
  tr_level<0x140, 8>(x, lane & 8);   // row_mirror
  tr_level<0x141, 4>(x, lane & 4);   // row_half_mirror
  tr_level<0x4e, 2>(x, lane & 2);    // quad_perm [2,3,0,1]
  tr_level<0xb1, 1>(x, lane & 1);    // quad_perm [1,0,3,2]
  return xor32_sum(xor16_sum(x[0]));
}

// Sums each of 16 vectors over the 16 lanes of every DPP row; lane l returns
// the row's total of vector x[l & 15] (the first four levels of
// transpose_reduce16, without the cross-row sums).
template <typename T>
__device__ inline T transpose_reduce16_row(T (&x)[16], int lane) {
  tr_level<0x140, 8>(x, lane & 8);   // row_mirror
  tr_level<0x141, 4>(x, lane & 4);   // row_half_mirror
  tr_level<0x4e, 2>(x, lane & 2);    // quad_perm [2,3,0,1]
  tr_level<0xb1, 1>(x, lane & 1);    // quad_perm [1,0,3,2]
  return x[0];
}

template <int CTRL, int ROW_MASK>
__device__ inline float dpp_add(float x) {
  return x + dpp_t<CTRL, ROW_MASK>(x);
}
template <int CTRL, int ROW_MASK>
__device__ inline double dpp_add(double x) {
  return x + dpp_t<CTRL, ROW_MASK>(x);
}

// sum over the 64 lanes, valid in lane 63
template <typename T>
__device__ inline T wave_sum_last(T x) {
  x = dpp_add<0xb1, 0xf>(x);   // quad_perm [1,0,3,2]
  x = dpp_add<0x4e, 0xf>(x);   // quad_perm [2,3,0,1]
  x = dpp_add<0x141, 0xf>(x);  // row_half_mirror
  x = dpp_add<0x140, 0xf>(x);  // row_mirror
  x = dpp_add<0x142, 0xa>(x);  // row_bcast:15 into rows 1, 3
  x = dpp_add<0x143, 0xc>(x);  // row_bcast:31 into rows 2, 3
  return x;
}

}  // namespace pscf
