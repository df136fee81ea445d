// B3LYP energy density and its analytic derivatives at one grid point.
//
// Functional as defined by the reference (pyscf_ipu/exchange_correlation/
// b3lyp.py:13-39 with lda.py:5, b88.py:13-29, lyp.py:13-45, vwn.py:13-54;
// formulas ported there from XCFun):
//   E = 0.08 Slater + 0.72 B88 (Slater + gradient term) + 0.19 VWN5 + 0.81 LYP,
// rho clipped to [1e-9, 1e12], |grad rho|^2 + eps.  The reference takes the
// derivatives by autodiff; here they are written out by hand:
//   VWN:  x = (3/(4 pi rho))^(1/6), G(x) = ln x + f/2 ln X - (1+f) ln(x-p0)
//         + pre atan(Q/(2x+p2)), E = rho p1 G, dE/drho = p1 (G - x G'/6),
//         G' = 1/x + f(2x+p2)/(2X) - (1+f)/(x-p0) - pre Q/(2X).
//   B88:  per-spin rho_a = rho/2, g_aa = gamma/4, chi^2 = g_aa rho_a^(-8/3),
//         D = 1 + 6 d chi asinh chi, D' = 6 d (asinh chi + chi/sqrt(1+chi^2)),
//         dE/drho_a = 4/3 c3 rho_a^(1/3) - 8/3 d rho_a^(1/3) chi^2 (chi D' - D)/D^2,
//         dE/dg_aa = -d (2D - chi D') / (D^2 rho_a^(4/3)).
//   LYP:  t = rho^(-1/3), P = 1/(1+Dd t), W = exp(-C t) P, delta = t (C + Dd P),
//         chain rule through dt/drho = -t/(3 rho).
#pragma once
#include <math.h>

namespace pscf {

// Elementary functions of the functional.  FP64: the libm functions.  FP32
// (device): the hardware's base-2 logarithm and exponential, reciprocal and
// square root (about 1 ulp each), with cbrt(x) = 2^(log2(x) / 3) and
// asinh(x) = ln(x + sqrt(x^2 + 1)) for x >= 0; relative errors stay near 1e-6,
// well inside what FP32 grid arithmetic gives anyway.
template <typename T>
struct XcMath {
  static __host__ __device__ inline T log_(T x) { return log(x); }
  static __host__ __device__ inline T exp_(T x) { return exp(x); }
  static __host__ __device__ inline T cbrt_(T x) { return cbrt(x); }
  static __host__ __device__ inline T sqrt_(T x) { return sqrt(x); }
  static __host__ __device__ inline T asinh_(T x) { return asinh(x); }
  static __host__ __device__ inline T div(T a, T b) { return a / b; }
};
#if defined(__HIP_DEVICE_COMPILE__)
template <>
struct XcMath<float> {
  static __device__ inline float log_(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
  static __device__ inline float exp_(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
  static __device__ inline float cbrt_(float x) { return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.0f / 3.0f)); }
  static __device__ inline float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ inline float asinh_(float x) { return log_(x + sqrt_(x * x + 1.0f)); }
  static __device__ inline float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
};
#endif

// Inputs of the functional: rho clipped to [1e-9, 1e12] and gamma = |grad rho|^2 + eps
// from rho and half its gradient (as the reference's exchange_correlation forms them).
template <typename T>
__host__ __device__ inline void b3lyp_inputs(T rho_in, T hx, T hy, T hz, T eps, T& rho, T& gam) {
  rho = rho_in < T(1e-9) ? T(1e-9) : (rho_in > T(1e12) ? T(1e12) : rho_in);
  gam = T(4) * (hx * hx + hy * hy + hz * hz) + eps;
}

// One of the three independent parts of the functional at (rho, gam) from
// b3lyp_inputs: K = 0 Slater + VWN5, K = 1 Becke 88, K = 2 LYP.  Each adds its
// energy per particle to exc, its dE/drho to vrho and its dE/dgamma part to
// vnorm; b3lyp_point is their sum, and a caller may evaluate the parts on
// different waves.
template <typename T, int K>
__host__ __device__ inline void b3lyp_part(T rho, T gam, T& exc, T& vrho, T& vnorm) {
  typedef XcMath<T> M;
  const T inv_rho = M::div(T(1), rho);
  if (K == 0) {
    // ---- Slater exchange, weight 0.08: e = -C rho^(1/3), v = 4/3 e
    const T r13 = M::cbrt_(rho);
    const T e_lda = T(-0.08 * 0.7385587663820229) * r13;
    const T v_lda = T(4.0 / 3.0) * e_lda;

    // ---- VWN5 correlation, weight 0.19
    const T p0 = T(-0.10498), p1 = T(0.0621813817393097900698817274255), p2 = T(3.72744), p3 = T(12.9352);
    const T fv = T(-0.10498 * 3.72744 / (0.10498 * 0.10498 - 0.10498 * 3.72744 + 12.9352) - 1.0);
    const T Qv = T(6.15199081975908);  // sqrt(4 p3 - p2^2)
    const T pre = T(3.72744 * (1.0 / 6.15199081975908 +
                               0.10498 / ((0.10498 * 0.10498 - 0.10498 * 3.72744 + 12.9352) * 6.15199081975908 /
                                          (3.72744 - 2.0 * 0.10498))));
    const T lnx = -M::log_(rho) * T(1.0 / 6.0) + T(-0.2387353263835302);  // ln(3/(4 pi))/6
    const T x = M::exp_(lnx);
    const T X = x * x + p2 * x + p3;
    const T G = lnx + T(0.5) * fv * M::log_(X) - (T(1) + fv) * M::log_(x - p0) + pre * atan(M::div(Qv, T(2) * x + p2));
    const T iX2 = M::div(T(1), T(2) * X);
    const T Gp = M::div(T(1), x) + fv * (T(2) * x + p2) * iX2 - M::div(T(1) + fv, x - p0) - pre * Qv * iX2;
    const T E_vwn = T(0.19) * rho * p1 * G;
    const T v_vwn = T(0.19) * p1 * (G - x * Gp * T(1.0 / 6.0));
    exc += e_lda + E_vwn * inv_rho;
    vrho += v_lda + v_vwn;
  } else if (K == 1) {
    // ---- Becke 88 exchange (incl. its Slater part), weight 0.72
    const T d = T(0.0042);
    const T c3 = T(-0.75 * 1.2407009817988000 * 2.0);  // -(3/4)(6/pi)^(1/3) * 2
    const T ra = T(0.5) * rho;
    const T ra13 = M::cbrt_(ra);
    const T ra43 = ra * ra13;
    const T gaa = T(0.25) * gam;
    const T chi2 = M::div(gaa, ra43 * ra43);
    const T chi = M::sqrt_(chi2);
    const T ash = M::asinh_(chi);
    const T Db = T(1) + T(6) * d * chi * ash;
    const T Dbp = T(6) * d * (ash + M::div(chi, M::sqrt_(T(1) + chi2)));
    const T iDb = M::div(T(1), Db);
    const T E_b88 = T(0.72) * (c3 * ra43 - T(2) * d * ra43 * chi2 * iDb);
    const T dE_dra = T(4.0 / 3.0) * c3 * ra13 - T(8.0 / 3.0) * d * ra13 * chi2 * (chi * Dbp - Db) * iDb * iDb;
    const T dE_dgaa = M::div(-d * (T(2) * Db - chi * Dbp) * iDb * iDb, ra43);
    exc += E_b88 * inv_rho;
    vrho += T(0.72 * 0.5) * dE_dra;
    vnorm += T(0.72 * 0.25) * dE_dgaa;
  } else {
    // ---- LYP correlation, weight 0.81
    const T A = T(0.04918), B = T(0.132), C = T(0.2533), Dd = T(0.349);
    const T CF = T(2.8712340001881915);  // 0.3 (3 pi^2)^(2/3)
    const T c0 = T(2.0);                 // 2^(11/3) (1/2)^(8/3)
    const T c1 = T((1.0 / 3.0 + 1.0 / 8.0) * 4.0);
    const T t = M::div(T(1), M::cbrt_(rho));
    const T P = M::div(T(1), T(1) + Dd * t);
    const T W = M::exp_(-C * t) * P;
    const T delta = t * (C + Dd * P);
    const T n53 = t * t * t * t * t;  // rho^(-5/3)
    const T K_ = (T(60) - T(14) * delta) * T(1.0 / 36.0) - c1;
    const T br = T(2) * CF * c0 * rho + gam * n53 * K_;
    const T AB4 = A * B * T(0.25);
    const T E_lyp_raw = -A * rho * P - AB4 * W * br;
    const T dt_drho = -t * inv_rho * T(1.0 / 3.0);
    const T dW_dt = -W * (C + Dd * P);
    const T ddelta_dt = C + Dd * P - t * Dd * Dd * P * P;
    const T dK_dt = T(-14.0 / 36.0) * ddelta_dt;
    const T dbr_drho = T(2) * CF * c0 + gam * (T(-5.0 / 3.0) * n53 * inv_rho * K_ + n53 * dK_dt * dt_drho);
    const T dE_lyp = -A * (P + Dd * P * P * t * T(1.0 / 3.0)) - AB4 * (dW_dt * dt_drho * br + W * dbr_drho);
    const T dEg_lyp = -AB4 * W * n53 * K_;
    exc += T(0.81) * E_lyp_raw * inv_rho;
    vrho += T(0.81) * dE_lyp;
    vnorm += T(0.81) * dEg_lyp;
  }
}

template <typename T>
__host__ __device__ inline void b3lyp_point(T rho_in, T hx, T hy, T hz, T eps, T& exc, T& vrho, T& vnorm) {
  T rho, gam;
  b3lyp_inputs(rho_in, hx, hy, hz, eps, rho, gam);
  exc = T(0);
  vrho = T(0);
  vnorm = T(0);
  b3lyp_part<T, 0>(rho, gam, exc, vrho, vnorm);
  b3lyp_part<T, 1>(rho, gam, exc, vrho, vnorm);
  b3lyp_part<T, 2>(rho, gam, exc, vrho, vnorm);
}

}  // namespace pscf
