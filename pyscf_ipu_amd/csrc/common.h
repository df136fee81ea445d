// Shared device helpers for the gfx950 kernels of libpyscf_mi355x.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pyscf_mi355x.h"

namespace pscf {

// ---- error plumbing ---------------------------------------------------------
void set_error(const char* fmt, ...);
int check_hip(hipError_t e, const char* what);
#define PSCF_HIP(call)                                              \
  do {                                                              \
    hipError_t _e = (call);                                         \
    if (_e != hipSuccess) return ::pscf::check_hip(_e, #call);      \
  } while (0)
// Raise a kernel's dynamic-LDS limit once per process (not a stream operation,
// kept out of graph-captured call sequences after the first call).
#define PSCF_SET_LDS(kernel, bytes)                                                                        \
  do {                                                                                                     \
    static int _pscf_lds_set = 0;                                                                          \
    if ((int)(bytes) > 65536 && (int)(bytes) > _pscf_lds_set) {                                            \
      PSCF_HIP(hipFuncSetAttribute((const void*)(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,     \
                                   (int)(bytes)));                                                         \
      _pscf_lds_set = (int)(bytes);                                                                        \
    }                                                                                                      \
  } while (0)

int device_max_lds();

#define PSCF_REQUIRE(cond, ...)                                     \
  do {                                                              \
    if (!(cond)) {                                                  \
      ::pscf::set_error(__VA_ARGS__);                               \
      return PSCF_EINVAL;                                           \
    }                                                               \
  } while (0)

// ---- libcint layout -----------------------------------------------------------
constexpr int ATM_SLOTS = 6;
constexpr int BAS_SLOTS = 8;
constexpr int MAX_PRIM = 8;
constexpr double PI = 3.14159265358979323846;
// Real solid harmonic factors libcint folds into s and p functions.
constexpr double ANG_S = 0.282094791773878143;
constexpr double ANG_P = 0.488602511902919921;

// ---- Boys function table --------------------------------------------------------
// F_m(T_i) for T_i = i * BOYS_H, i < BOYS_NT, m < BOYS_NM, row-major [i][m].
// Evaluated as a Taylor series about the nearest node,
//   F_m(T_i + d) = sum_k F_{m+k}(T_i) (-d)^k / k!,
// and beyond BOYS_TMAX by the asymptote F_0 = sqrt(pi/T)/2 with upward recursion.
constexpr int BOYS_NT = 400;
constexpr double BOYS_H = 0.1;
constexpr double BOYS_TMAX = 39.9;
constexpr int BOYS_NM = 15;
constexpr int BOYS_TAYLOR = 8;  // terms beyond the leading one (FP64)

template <typename T>
__device__ inline T dexp(T x);
template <>
__device__ inline double dexp<double>(double x) { return exp(x); }
template <>
__device__ inline float dexp<float>(float x) { return expf(x); }

// 1 / x: the hardware reciprocal in FP32 (1 ulp), a division in FP64
template <typename T>
__device__ inline T drcp(T x);
template <>
__device__ inline double drcp<double>(double x) { return 1.0 / x; }
template <>
__device__ inline float drcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }

template <typename T>
__device__ inline T dsqrt(T x);
template <>
__device__ inline double dsqrt<double>(double x) { return sqrt(x); }
template <>
__device__ inline float dsqrt<float>(float x) { return sqrtf(x); }

// F_0..F_mmax at T from a table staged in LDS (`tab`, row-major [BOYS_NT][BOYS_NM]).
// NTERMS Taylor terms after the first for F_mmax, then downward recursion.
template <typename T, int MMAX, int NTERMS>
__device__ inline void boys(T t, const T* __restrict__ tab, T* f) {
  if (t < T(BOYS_TMAX)) {
    int i = (int)(t * T(1.0 / BOYS_H) + T(0.5));
    T d = T(i) * T(BOYS_H) - t;  // (-delta)
    const T* row = tab + i * BOYS_NM + MMAX;
    T acc = row[NTERMS];
#pragma unroll
    for (int k = NTERMS; k > 0; --k) acc = row[k - 1] + acc * d * T(1.0 / k);
    f[MMAX] = acc;
    if (MMAX > 0) {
      T e = dexp<T>(-t);
#pragma unroll
      for (int m = MMAX; m > 0; --m) f[m - 1] = (T(2) * t * f[m] + e) * T(1.0 / (2 * m - 1));
    }
  } else {
    T inv = T(1) / t;
    f[0] = T(0.5) * dsqrt<T>(T(PI) * inv);
#pragma unroll
    for (int m = 0; m < MMAX; ++m) f[m + 1] = f[m] * T(2 * m + 1) * T(0.5) * inv;
  }
}

// Rys rule (u_r = t_r^2, w_r) of n points for the weight exp(-T u)/(2 sqrt u) on
// [0,1] from its moments mu_k = F_k(T), k < 2n: the monic orthogonal polynomial's
// coefficients solve the Hankel system, its roots are found in closed form
// (n <= 3) and the weights solve the Vandermonde system sum_r w_r u_r^k = mu_k.
// Computed in FP64.
template <int N>
__device__ inline void rys_from_moments(const double* mu, double* u, double* w);

template <>
__device__ inline void rys_from_moments<1>(const double* mu, double* u, double* w) {
  u[0] = mu[1] / mu[0];
  w[0] = mu[0];
}

template <>
__device__ inline void rys_from_moments<2>(const double* mu, double* u, double* w) {
  double det = mu[0] * mu[2] - mu[1] * mu[1];
  double a = (mu[1] * mu[2] - mu[0] * mu[3]) / det;  // p(u) = u^2 + a u + b
  double b = (mu[1] * mu[3] - mu[2] * mu[2]) / det;
  double disc = sqrt(fmax(a * a - 4.0 * b, 0.0));
  // numerically stable pair of roots
  double qq = -0.5 * (a + (a >= 0 ? disc : -disc));
  double r1 = qq, r2 = b / qq;
  double lo = fmin(r1, r2), hi = fmax(r1, r2);
  u[0] = lo;
  u[1] = hi;
  w[0] = (mu[1] - hi * mu[0]) / (lo - hi);
  w[1] = mu[0] - w[0];
}

template <>
__device__ inline void rys_from_moments<3>(const double* mu, double* u, double* w) {
  // Hankel system H c = -(mu3, mu4, mu5), p(u) = u^3 + c2 u^2 + c1 u + c0.
  double h00 = mu[0], h01 = mu[1], h02 = mu[2], h11 = mu[2], h12 = mu[3], h22 = mu[4];
  double r0 = -mu[3], r1 = -mu[4], r2 = -mu[5];
  // Cholesky of the (positive definite) Hankel matrix.
  double l00 = sqrt(h00);
  double l10 = h01 / l00, l20 = h02 / l00;
  double l11 = sqrt(h11 - l10 * l10);
  double l21 = (h12 - l20 * l10) / l11;
  double l22 = sqrt(h22 - l20 * l20 - l21 * l21);
  double y0 = r0 / l00;
  double y1 = (r1 - l10 * y0) / l11;
  double y2 = (r2 - l20 * y0 - l21 * y1) / l22;
  double c2 = y2 / l22;
  double c1 = (y1 - l21 * c2) / l11;
  double c0 = (y0 - l10 * c1 - l20 * c2) / l00;
  // Depressed cubic x^3 + p x + q with u = x - c2/3; three real roots.
  double s = c2 * (1.0 / 3.0);
  double p = c1 - c2 * s;
  double q = 2.0 * s * s * s - s * c1 + c0;
  double m = 2.0 * sqrt(fmax(-p * (1.0 / 3.0), 0.0));
  double arg = (m > 0) ? 3.0 * q / (p * m) : 0.0;
  arg = fmin(1.0, fmax(-1.0, arg));
  double th = acos(arg) * (1.0 / 3.0);
  double x0 = m * cos(th), x1 = m * cos(th - 2.0943951023931954923), x2 = m * cos(th - 4.1887902047863909846);
  double rr[3] = {x0 - s, x1 - s, x2 - s};
  // Newton polish on p(u).
#pragma unroll
  for (int k = 0; k < 3; ++k) {
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      double x = rr[k];
      double f = ((x + c2) * x + c1) * x + c0;
      double df = (3.0 * x + 2.0 * c2) * x + c1;
      rr[k] = x - f / df;
    }
  }
  // sort ascending
  double a0 = fmin(rr[0], fmin(rr[1], rr[2])), a2 = fmax(rr[0], fmax(rr[1], rr[2]));
  double a1 = rr[0] + rr[1] + rr[2] - a0 - a2;
  u[0] = a0;
  u[1] = a1;
  u[2] = a2;
  w[0] = (mu[2] - (a1 + a2) * mu[1] + a1 * a2 * mu[0]) / ((a0 - a1) * (a0 - a2));
  w[1] = (mu[2] - (a0 + a2) * mu[1] + a0 * a2 * mu[0]) / ((a1 - a0) * (a1 - a2));
  w[2] = mu[0] - w[0] - w[1];
}

// Stage the Boys table from global memory into LDS in the precision T.
template <typename T>
__device__ inline void stage_boys(T* lds, const double* __restrict__ g) {
  for (int i = threadIdx.x; i < BOYS_NT * BOYS_NM; i += blockDim.x) lds[i] = (T)g[i];
  __syncthreads();
}

__host__ __device__ inline int64_t tri(int64_t a) { return a * (a + 1) / 2; }

}  // namespace pscf
