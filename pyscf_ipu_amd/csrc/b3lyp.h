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

template <typename T>
__host__ __device__ inline void b3lyp_point(T rho_in, T hx, T hy, T hz, T eps, T& exc, T& vrho, T& vnorm) {
  // inputs: rho and half its gradient (as the reference's exchange_correlation forms them)
  const T rho = rho_in < T(1e-9) ? T(1e-9) : (rho_in > T(1e12) ? T(1e12) : rho_in);
  const T gam = T(4) * (hx * hx + hy * hy + hz * hz) + eps;
  const T r13 = cbrt(rho);
  const T inv_rho = T(1) / rho;

  // ---- Slater exchange, weight 0.08: e = -C rho^(1/3), v = 4/3 e
  const T e_lda = T(-0.08 * 0.7385587663820229) * r13;
  const T v_lda = T(4.0 / 3.0) * e_lda;

  // ---- VWN5 correlation, weight 0.19
  const T p0 = T(-0.10498), p1 = T(0.0621813817393097900698817274255), p2 = T(3.72744), p3 = T(12.9352);
  const T fv = T(-0.10498 * 3.72744 / (0.10498 * 0.10498 - 0.10498 * 3.72744 + 12.9352) - 1.0);
  const T Qv = T(6.15199081975908);  // sqrt(4 p3 - p2^2)
  const T pre = T(3.72744 * (1.0 / 6.15199081975908 +
                             0.10498 / ((0.10498 * 0.10498 - 0.10498 * 3.72744 + 12.9352) * 6.15199081975908 /
                                        (3.72744 - 2.0 * 0.10498))));
  const T lnx = -log(rho) * T(1.0 / 6.0) + T(-0.2387353263835302);  // ln(3/(4 pi))/6
  const T x = exp(lnx);
  const T X = x * x + p2 * x + p3;
  const T G = lnx + T(0.5) * fv * log(X) - (T(1) + fv) * log(x - p0) + pre * atan(Qv / (T(2) * x + p2));
  const T Gp = T(1) / x + fv * (T(2) * x + p2) / (T(2) * X) - (T(1) + fv) / (x - p0) - pre * Qv / (T(2) * X);
  const T E_vwn = T(0.19) * rho * p1 * G;
  const T v_vwn = T(0.19) * p1 * (G - x * Gp * T(1.0 / 6.0));

  // ---- Becke 88 exchange (incl. its Slater part), weight 0.72
  const T d = T(0.0042);
  const T c3 = T(-0.75 * 1.2407009817988000 * 2.0);  // -(3/4)(6/pi)^(1/3) * 2
  const T ra = T(0.5) * rho;
  const T ra13 = cbrt(ra);
  const T ra43 = ra * ra13;
  const T gaa = T(0.25) * gam;
  const T chi2 = gaa / (ra43 * ra43);
  const T chi = sqrt(chi2);
  const T ash = asinh(chi);
  const T Db = T(1) + T(6) * d * chi * ash;
  const T Dbp = T(6) * d * (ash + chi / sqrt(T(1) + chi2));
  const T E_b88 = T(0.72) * (c3 * ra43 - T(2) * d * ra43 * chi2 / Db);
  const T dE_dra = T(4.0 / 3.0) * c3 * ra13 - T(8.0 / 3.0) * d * ra13 * chi2 * (chi * Dbp - Db) / (Db * Db);
  const T dE_dgaa = -d * (T(2) * Db - chi * Dbp) / (Db * Db * ra43);
  const T v_b88 = T(0.72 * 0.5) * dE_dra;
  const T vn_b88 = T(0.72 * 0.25) * dE_dgaa;

  // ---- LYP correlation, weight 0.81
  const T A = T(0.04918), B = T(0.132), C = T(0.2533), Dd = T(0.349);
  const T CF = T(2.8712340001881915);  // 0.3 (3 pi^2)^(2/3)
  const T c0 = T(2.0);                 // 2^(11/3) (1/2)^(8/3)
  const T c1 = T((1.0 / 3.0 + 1.0 / 8.0) * 4.0);
  const T t = T(1) / r13;
  const T P = T(1) / (T(1) + Dd * t);
  const T W = exp(-C * t) * P;
  const T delta = t * (C + Dd * P);
  const T n53 = t * t * t * t * t;  // rho^(-5/3)
  const T K = (T(60) - T(14) * delta) * T(1.0 / 36.0) - c1;
  const T br = T(2) * CF * c0 * rho + gam * n53 * K;
  const T AB4 = A * B * T(0.25);
  const T E_lyp_raw = -A * rho * P - AB4 * W * br;
  const T dt_drho = -t * inv_rho * T(1.0 / 3.0);
  const T dW_dt = -W * (C + Dd * P);
  const T ddelta_dt = C + Dd * P - t * Dd * Dd * P * P;
  const T dK_dt = T(-14.0 / 36.0) * ddelta_dt;
  const T dbr_drho = T(2) * CF * c0 + gam * (T(-5.0 / 3.0) * n53 * inv_rho * K + n53 * dK_dt * dt_drho);
  const T dE_lyp = -A * (P + Dd * P * P * t * T(1.0 / 3.0)) - AB4 * (dW_dt * dt_drho * br + W * dbr_drho);
  const T dEg_lyp = -AB4 * W * n53 * K;
  const T E_lyp = T(0.81) * E_lyp_raw;

  exc = e_lda + (E_vwn + E_b88 + E_lyp) * inv_rho;
  vrho = v_lda + v_vwn + v_b88 + T(0.81) * dE_lyp;
  vnorm = vn_b88 + T(0.81) * dEg_lyp;
}

}  // namespace pscf
