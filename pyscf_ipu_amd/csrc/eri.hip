// Two-electron repulsion integrals and one-electron integrals on gfx950.
//
// The reference's ERI generator is libcint's Rys-quadrature scheme
// (pyscf_ipu/electron_repulsion/int2e_sph.cpp: CINTg0_2e :1785-1900, 2D
// recurrence CINTg0_2e_2d :1475-1600, primitive loop CINT2e_loop_nopt
// :2452-2723, one shell quartet per IPU tile).  Here the same quadrature runs
// one contracted shell quartet per GPU lane, with every wavefront of a launch
// in a single angular class (ss|ss), (ps|ss), (ps|ps), (pp|ss), (pp|ps) or
// (pp|pp), so the per-class code is fully unrolled and branch-free:
//   * Boys F_m(T) from a Taylor table staged in LDS (F_0..F_14 on a 0.1 grid
//     up to T = 39.9, asymptote above),
//   * Rys roots/weights from the Boys moments in closed form (1-3 roots), FP64,
//   * 2D recurrence G(n, m) per root and axis, horizontal transfer to
//     (i j | k l), accumulation of sum_r w_r Ix Iy Iz over primitive quartets.
// Launches for the packed s8 array enumerate (bra pair, ket pair) directly from
// a class-sorted shell-pair list, so no quartet list is materialised.
#include <cmath>
#include <mutex>
#include <vector>

#include "batch.h"
#include "common.h"
#include "wave.h"

namespace pscf {

// --------------------------------------------------------------- Boys table ---
__device__ double g_boys[BOYS_NT * BOYS_NM];
static std::mutex g_boys_mu;
static bool g_boys_ready[64];

static void boys_row_host(long double t, long double* f) {
  const int mmax = BOYS_NM - 1;
  long double e = expl(-t);
  if (t < 12.0L) {
    long double term = 1.0L / (2 * mmax + 1), sum = term;
    for (int k = 1; k < 400; k++) {
      term *= 2.0L * t / (2 * mmax + 2 * k + 1);
      sum += term;
      if (term < 1e-22L * sum) break;
    }
    f[mmax] = e * sum;
    for (int m = mmax; m > 0; m--) f[m - 1] = (2.0L * t * f[m] + e) / (2 * m - 1);
  } else {
    // Upward recursion loses digits for m ~ T; use the series from a high m
    // started by a long continued sum, then recurse downward.
    const int mtop = mmax + 60;
    long double term = 1.0L / (2 * mtop + 1), sum = term;
    for (int k = 1; k < 4000; k++) {
      term *= 2.0L * t / (2 * mtop + 2 * k + 1);
      sum += term;
      if (term < 1e-24L * sum) break;
    }
    long double fm = e * sum;
    for (int m = mtop; m > 0; m--) {
      fm = (2.0L * t * fm + e) / (2 * m - 1);
      if (m - 1 <= mmax) f[m - 1] = fm;
    }
  }
}

static int ensure_boys_table() {
  int dev = 0;
  PSCF_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_boys_mu);
  if (dev < 64 && g_boys_ready[dev]) return PSCF_OK;
  std::vector<double> tab(BOYS_NT * BOYS_NM);
  long double f[BOYS_NM];
  for (int i = 0; i < BOYS_NT; i++) {
    boys_row_host((long double)i * (long double)BOYS_H, f);
    for (int m = 0; m < BOYS_NM; m++) tab[i * BOYS_NM + m] = (double)f[m];
  }
  PSCF_HIP(hipMemcpyToSymbol(HIP_SYMBOL(g_boys), tab.data(), sizeof(double) * tab.size(), 0, hipMemcpyHostToDevice));
  if (dev < 64) g_boys_ready[dev] = true;
  return PSCF_OK;
}

// ------------------------------------------------------------------- shells ---
struct ShellD {
  double r[3];
  const double* e;
  const double* c;
  int np;
  int ao0;
};

__device__ inline ShellD load_shell(const int* __restrict__ atm, const int* __restrict__ bas,
                                    const double* __restrict__ env, const int* __restrict__ ao_loc, int ib) {
  ShellD s;
  const int* b = bas + ib * BAS_SLOTS;
  const double* r = env + atm[b[0] * ATM_SLOTS + 1];
  s.r[0] = r[0];
  s.r[1] = r[1];
  s.r[2] = r[2];
  s.np = b[2];
  s.e = env + b[5];
  s.c = env + b[6];
  s.ao0 = ao_loc[ib];
  return s;
}

// Prefix sum of shell widths (2l+1) into LDS.
__device__ inline void stage_ao_loc(int* ao_loc, const int* __restrict__ bas, int nbas) {
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int i = 0; i < nbas; i++) {
      ao_loc[i] = acc;
      acc += 2 * bas[i * BAS_SLOTS + 1] + 1;
    }
    ao_loc[nbas] = acc;
  }
}

template <int L>
struct NComp {
  static constexpr int v = L ? 3 : 1;
};

// Cartesian exponent of component f (x, y, z for p) along axis d.
template <int L>
__device__ constexpr int cexp(int f, int d) {
  return L ? (f == d ? 1 : 0) : 0;
}

// ------------------------------------------------------- shell-quartet ERIs ---
// acc[((fi*NJ + fj)*NK + fk)*NL + fl] += (ij|kl), contracted, spherical (l <= 1).
// T: arithmetic of the recurrence/accumulation; BT: Boys-table precision.
// NPC > 0: every shell has NPC primitives (STO-3G: 3), so the primitive loops
// unroll and the ket-pair data (exponent sum, its reciprocal, the Gaussian
// product prefactor with its exp) is computed once per contracted quartet
// instead of once per bra primitive pair; NPC = 0 reads the counts at run time.
// SCREEN (FP32, quartets of four 3-primitive shells): bra primitive pairs with
// a negligible Gaussian-product prefactor skip their ket loop (bound below).
template <typename T, typename BT, int NTERMS, int LI, int LJ, int LK, int LL, int NPC, bool SCREEN>
__device__ inline void shell_quartet_np(const ShellD& A, const ShellD& B, const ShellD& C, const ShellD& D,
                                        const BT* __restrict__ tab, T* acc, bool screen) {
  constexpr int NR = (LI + LJ + LK + LL) / 2 + 1;
  constexpr int NMAX = LI + LJ, MMAX = LK + LL;
  constexpr int NI = NComp<LI>::v, NJ = NComp<LJ>::v, NK = NComp<LK>::v, NL = NComp<LL>::v;
  constexpr int NC = NI * NJ * NK * NL;
  constexpr bool SSSS = LI + LJ + LK + LL == 0;  // value = prefactor * F_0: no root needed
  constexpr int NKP = NPC > 0 ? NPC * NPC : 1;
  const T angc = (T)((LI ? ANG_P : ANG_S) * (LJ ? ANG_P : ANG_S) * (LK ? ANG_P : ANG_S) * (LL ? ANG_P : ANG_S));
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = T(0);
  T AB[3], CD[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    AB[d] = (T)(A.r[d] - B.r[d]);
    CD[d] = (T)(C.r[d] - D.r[d]);
  }
  const T rab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
  const T rcd2 = CD[0] * CD[0] + CD[1] * CD[1] + CD[2] * CD[2];
  // ket pairs, hoisted (NPC > 0)
  T k_akl[NKP], k_r[NKP], k_al[NKP], k_kcd[NKP];
  if constexpr (NPC > 0) {
#pragma unroll
    for (int pc = 0; pc < NPC; ++pc)
#pragma unroll
      for (int pd = 0; pd < NPC; ++pd) {
        const int q = pc * NPC + pd;
        const T ak = (T)C.e[pc], al = (T)D.e[pd];
        k_akl[q] = ak + al;
        k_r[q] = drcp<T>(k_akl[q]);
        k_al[q] = al * k_r[q];
        k_kcd[q] = dexp<T>(-ak * k_al[q] * rcd2) * (T)C.c[pc] * (T)D.c[pd];
      }
  }
  const int npa = NPC > 0 ? NPC : A.np, npb = NPC > 0 ? NPC : B.np;
  const int npc = NPC > 0 ? NPC : C.np, npd = NPC > 0 ? NPC : D.np;
  // bra primitive pairs in a rolled loop (code size); ket pairs unrolled when
  // NPC > 0, so the hoisted arrays are indexed at compile time
#pragma unroll 1
  for (int pa = 0; pa < npa; ++pa) {
    const T ai = (T)A.e[pa], ca = (T)A.c[pa];
#pragma unroll 1
    for (int pb = 0; pb < npb; ++pb) {
      const T aj = (T)B.e[pb];
      const T aij = ai + aj, r_aij = drcp<T>(aij);
      const T kab = dexp<T>(-ai * aj * r_aij * rab2) * ca * (T)B.c[pb];
      // SCREEN: a bra primitive pair whose Gaussian-product prefactor is below
      // the bound skips its ket loop.  What it would add to an ERI is at most
      // 34.99 r_aij |kab| (akl^-3/2 <= 5.1) (sum of |kcd| over the 3 x 3 ket
      // primitive pairs, <= 9 x 15^2 for STO-3G shells) angc (<= 0.06) times
      // the 1D-integral products (< 1e5 for molecules under 20 bohr): below
      // 1e-10, far under the FP32 rounding of the ERIs themselves (~1e-7 of the
      // largest).  The lanes of a wave mostly share their bra pair, so whole
      // waves skip together.
      if constexpr (SCREEN && sizeof(T) == 4)
        if (screen && fabsf(kab) * r_aij < 4e-20f) continue;
      T PA[3], P[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        PA[d] = -aj * r_aij * AB[d];  // P - A
        P[d] = (T)A.r[d] + PA[d];
      }
#pragma unroll
      for (int pc = 0; pc < npc; ++pc) {
#pragma unroll
        for (int pd = 0; pd < npd; ++pd) {
          T akl, r_akl, alr, kcd;
          if constexpr (NPC > 0) {
            const int q = pc * NPC + pd;
            akl = k_akl[q];
            r_akl = k_r[q];
            alr = k_al[q];
            kcd = k_kcd[q];
          } else {
            const T ak = (T)C.e[pc], al = (T)D.e[pd];
            akl = ak + al;
            r_akl = drcp<T>(akl);
            alr = al * r_akl;
            kcd = dexp<T>(-ak * alr * rcd2) * (T)C.c[pc] * (T)D.c[pd];
          }
          T QC[3], PQ[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            QC[d] = -alr * CD[d];  // Q - C
            PQ[d] = P[d] - ((T)C.r[d] + QC[d]);
          }
          const T r_s = drcp<T>(aij + akl);
          const T rho = aij * akl * r_s;
          const T tt = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
          // 2 pi^(5/2) / (aij akl sqrt(aij + akl)) * K_ab * K_cd * angular factors
          const T pref = T(34.986836655249725) * r_aij * r_akl * dsqrt<T>(r_s) * kab * kcd * angc;
          if constexpr (SSSS) {
            BT f[1];
            boys<BT, 0, NTERMS>((BT)tt, tab, f);
            acc[0] += pref * (T)f[0];
            continue;
          }
          BT f[2 * NR];
          boys<BT, 2 * NR - 1, NTERMS>((BT)tt, tab, f);
          T u[NR], w[NR];
          if constexpr (NR == 1) {
            u[0] = (T)(f[1] / f[0]);
            w[0] = (T)f[0];
          } else {
            double mu[2 * NR], ud[NR], wd[NR];
#pragma unroll
            for (int k = 0; k < 2 * NR; ++k) mu[k] = (double)f[k];
            rys_from_moments<NR>(mu, ud, wd);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
              u[r] = (T)ud[r];
              w[r] = (T)wd[r];
            }
          }
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const T t2 = u[r];
            const T b00 = T(0.5) * t2 * r_s;
            const T b10 = T(0.5) * r_aij * (T(1) - akl * t2 * r_s);
            const T b01 = T(0.5) * r_akl * (T(1) - aij * t2 * r_s);
            T I[3][LI + 1][LJ + 1][LK + 1][LL + 1];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              const T c00 = PA[d] - akl * PQ[d] * t2 * r_s;
              const T c0p = QC[d] + aij * PQ[d] * t2 * r_s;
              T g[NMAX + 2][MMAX + 2];
              g[0][0] = T(1);
#pragma unroll
              for (int n = 0; n < NMAX; ++n) g[n + 1][0] = c00 * g[n][0] + (n > 0 ? T(n) * b10 * g[n - 1][0] : T(0));
#pragma unroll
              for (int m = 0; m < MMAX; ++m)
#pragma unroll
                for (int n = 0; n <= NMAX; ++n)
                  g[n][m + 1] = c0p * g[n][m] + (m > 0 ? T(m) * b01 * g[n][m - 1] : T(0)) +
                                (n > 0 ? T(n) * b00 * g[n - 1][m] : T(0));
#pragma unroll
              for (int i = 0; i <= LI; ++i)
#pragma unroll
                for (int k = 0; k <= LK; ++k) {
                  I[d][i][0][k][0] = g[i][k];
                  if constexpr (LJ > 0) I[d][i][1][k][0] = g[i + 1][k] + AB[d] * g[i][k];
                  if constexpr (LL > 0) I[d][i][0][k][1] = g[i][k + 1] + CD[d] * g[i][k];
                  if constexpr (LJ > 0 && LL > 0)
                    I[d][i][1][k][1] =
                        g[i + 1][k + 1] + AB[d] * g[i][k + 1] + CD[d] * (g[i + 1][k] + AB[d] * g[i][k]);
                }
            }
            const T wr = w[r] * pref;
#pragma unroll
            for (int fi = 0; fi < NI; ++fi)
#pragma unroll
              for (int fj = 0; fj < NJ; ++fj)
#pragma unroll
                for (int fk = 0; fk < NK; ++fk)
#pragma unroll
                  for (int fl = 0; fl < NL; ++fl) {
                    T v = wr;
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                      v *= I[d][cexp<LI>(fi, d)][cexp<LJ>(fj, d)][cexp<LK>(fk, d)][cexp<LL>(fl, d)];
                    acc[((fi * NJ + fj) * NK + fk) * NL + fl] += v;
                  }
          }
        }
      }
    }
  }
}

// Dispatch on the primitive counts: in FP32 and the one-root classes ((ss|ss),
// (ps|ss), the most numerous) four shells with 3 primitives (every STO-3G
// shell) take the unrolled, ket-hoisted form; the multi-root classes and FP64
// keep rolled loops (unrolled, they need more registers than a lane has).
// SCREEN (the batched STO-3G form) enables the primitive-pair screening of
// shell_quartet_np on quartets of four 3-primitive shells.
template <typename T, typename BT, int NTERMS, int LI, int LJ, int LK, int LL, bool SCREEN = false>
__device__ inline void shell_quartet(const ShellD& A, const ShellD& B, const ShellD& C, const ShellD& D,
                                     const BT* __restrict__ tab, T* acc) {
  constexpr bool ONE_ROOT = LI + LJ + LK + LL <= 1 && sizeof(T) == 4;  // FP32 (FP64 would need 2x the registers)
  const bool np3 = A.np == 3 && B.np == 3 && C.np == 3 && D.np == 3;
  if (ONE_ROOT && np3)
    shell_quartet_np<T, BT, NTERMS, LI, LJ, LK, LL, ONE_ROOT ? 3 : 0, SCREEN>(A, B, C, D, tab, acc, true);
  else
    shell_quartet_np<T, BT, NTERMS, LI, LJ, LK, LL, 0, SCREEN>(A, B, C, D, tab, acc, np3);
}

template <typename BT>
struct BoysTerms;
template <>
struct BoysTerms<double> {
  static constexpr int v = BOYS_TAYLOR;
};
template <>
struct BoysTerms<float> {
  static constexpr int v = 5;
};

// Pair class -> (l of first shell, l of second shell): 0 = ss, 1 = ps, 2 = pp.
template <int X>
struct PairL {
  static constexpr int a = X >= 1 ? 1 : 0;
  static constexpr int b = X == 2 ? 1 : 0;
};

// Boys precision per class: roots from >= 2 moments are computed from FP64 Boys
// values even in the FP32 kernels (Hankel systems amplify input error).
template <typename T, int NR>
struct BoysType {
  using type = double;
};
template <>
struct BoysType<float, 1> {
  using type = float;
};

// Quartet q of class (X, Y) of a molecule's s8 array: pair a of class X and
// pair b of class Y (a >= b within one class), pairs listed from x0 / y0.
template <typename T, int X, int Y>
__device__ inline void s8_quartet(const int* __restrict__ atm, const int* __restrict__ bas,
                                  const double* __restrict__ env, const int* __restrict__ ao_loc,
                                  const int* __restrict__ pairs, int x0, int nx, int y0, int ny, int64_t q,
                                  const double* tab64, const float* tab32, T* __restrict__ out) {
  constexpr int LI = PairL<X>::a, LJ = PairL<X>::b, LK = PairL<Y>::a, LL = PairL<Y>::b;
  constexpr int NR = (LI + LJ + LK + LL) / 2 + 1;
  using BT = typename BoysType<T, NR>::type;
  const BT* tab;
  if constexpr (sizeof(BT) == 8) tab = tab64; else tab = tab32;
  const bool tri_launch = (X == Y);
  const int64_t total = tri_launch ? tri((int64_t)nx) : (int64_t)nx * ny;
  if (q >= total) return;
  int64_t a, b;
  if (tri_launch) {
    a = (int64_t)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
    while (tri(a + 1) <= q) ++a;
    while (tri(a) > q) --a;
    b = q - tri(a);
  } else {
    a = q / ny;
    b = q - a * ny;
  }
  const int i = pairs[2 * (x0 + a)], j = pairs[2 * (x0 + a) + 1];
  const int k = pairs[2 * (y0 + b)], l = pairs[2 * (y0 + b) + 1];
  const ShellD A = load_shell(atm, bas, env, ao_loc, i), B = load_shell(atm, bas, env, ao_loc, j);
  const ShellD C = load_shell(atm, bas, env, ao_loc, k), D = load_shell(atm, bas, env, ao_loc, l);
  constexpr int NI = NComp<LI>::v, NJ = NComp<LJ>::v, NK = NComp<LK>::v, NL = NComp<LL>::v;
  T acc[NI * NJ * NK * NL];
  shell_quartet<T, BT, BoysTerms<BT>::v, LI, LJ, LK, LL>(A, B, C, D, tab, acc);
  const bool diag_bra = (i == j), diag_ket = (k == l), same_pair = tri_launch && (a == b);
#pragma unroll
  for (int fi = 0; fi < NI; ++fi)
#pragma unroll
    for (int fj = 0; fj < NJ; ++fj)
#pragma unroll
      for (int fk = 0; fk < NK; ++fk)
#pragma unroll
        for (int fl = 0; fl < NL; ++fl) {
          const int64_t p = A.ao0 + fi, qq = B.ao0 + fj, r = C.ao0 + fk, s = D.ao0 + fl;
          if (diag_bra && qq > p) continue;
          if (diag_ket && s > r) continue;
          const int64_t pq = p >= qq ? tri(p) + qq : tri(qq) + p;
          const int64_t rs = r >= s ? tri(r) + s : tri(s) + r;
          if (same_pair && rs > pq) continue;
          const int64_t idx = pq >= rs ? tri(pq) + rs : tri(rs) + pq;
          out[idx] = acc[((fi * NJ + fj) * NK + fk) * NL + fl];
        }
}

// The six shell-pair classes of one molecule in one launch, so they run side
// by side rather than one after another: blocks [b0[c], b0[c + 1]) take class
// S8_ORDER[c], the heaviest first.  off / cnt: each pair class's first pair
// and count in `pairs` (s-s, p-s, p-p).
constexpr int S8_CLASSES = 6;
struct S8Launch {
  int off[3], cnt[3];
  int b0[S8_CLASSES + 1];
};

template <typename T>
__global__ void __launch_bounds__(256) eri_s8_kernel(const int* __restrict__ atm, const int* __restrict__ bas,
                                                     const double* __restrict__ env, int nbas,
                                                     const int* __restrict__ pairs, S8Launch L,
                                                     T* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* tab64 = reinterpret_cast<double*>(smem);
  float* tab32 = reinterpret_cast<float*>(smem + sizeof(double) * BOYS_NT * BOYS_NM);
  int* ao_loc = reinterpret_cast<int*>(smem + (sizeof(double) + sizeof(float)) * BOYS_NT * BOYS_NM);
  for (int i = threadIdx.x; i < BOYS_NT * BOYS_NM; i += blockDim.x) {
    tab64[i] = g_boys[i];
    if constexpr (sizeof(T) == 4) tab32[i] = (float)g_boys[i];
  }
  stage_ao_loc(ao_loc, bas, nbas);
  __syncthreads();
  const int b = blockIdx.x;
  const int c = b < L.b0[1] ? 0 : b < L.b0[2] ? 1 : b < L.b0[3] ? 2 : b < L.b0[4] ? 3 : b < L.b0[5] ? 4 : 5;
  const int64_t q = (int64_t)(b - L.b0[c]) * blockDim.x + threadIdx.x;
  const int* o = L.off;
  const int* n = L.cnt;
  switch (c) {  // S8_ORDER: (2,2), (2,1), (2,0), (1,1), (1,0), (0,0)
    case 0: s8_quartet<T, 2, 2>(atm, bas, env, ao_loc, pairs, o[2], n[2], o[2], n[2], q, tab64, tab32, out); break;
    case 1: s8_quartet<T, 2, 1>(atm, bas, env, ao_loc, pairs, o[2], n[2], o[1], n[1], q, tab64, tab32, out); break;
    case 2: s8_quartet<T, 2, 0>(atm, bas, env, ao_loc, pairs, o[2], n[2], o[0], n[0], q, tab64, tab32, out); break;
    case 3: s8_quartet<T, 1, 1>(atm, bas, env, ao_loc, pairs, o[1], n[1], o[1], n[1], q, tab64, tab32, out); break;
    case 4: s8_quartet<T, 1, 0>(atm, bas, env, ao_loc, pairs, o[1], n[1], o[0], n[0], q, tab64, tab32, out); break;
    default: s8_quartet<T, 0, 0>(atm, bas, env, ao_loc, pairs, o[0], n[0], o[0], n[0], q, tab64, tab32, out); break;
  }
}

// Quartets of an explicit list; each lane canonicalises its quartet to a class
// with l_i >= l_j, l_k >= l_l, (l_i + l_j) >= (l_k + l_l) and writes the block
// back in the caller's (i, j, k, l) column-major order.
template <typename T, int X, int Y>
__device__ inline void quartet_list_class(const ShellD* S, bool swap_ij, bool swap_kl, bool swap_braket,
                                          const double* tab64, const float* tab32, T* __restrict__ dst) {
  constexpr int LI = PairL<X>::a, LJ = PairL<X>::b, LK = PairL<Y>::a, LL = PairL<Y>::b;
  constexpr int NR = (LI + LJ + LK + LL) / 2 + 1;
  using BT = typename BoysType<T, NR>::type;
  const BT* tab;
  if constexpr (sizeof(BT) == 8) tab = tab64; else tab = tab32;
  constexpr int NI = NComp<LI>::v, NJ = NComp<LJ>::v, NK = NComp<LK>::v, NL = NComp<LL>::v;
  T acc[NI * NJ * NK * NL];
  shell_quartet<T, BT, BoysTerms<BT>::v, LI, LJ, LK, LL>(S[0], S[1], S[2], S[3], tab, acc);
  // canonical component indices -> caller's (a, b, c, d)
  int dims[4] = {NI, NJ, NK, NL};
  // caller order: positions of canonical slots
  int perm[4] = {0, 1, 2, 3};  // perm[caller slot] = canonical slot
  if (swap_braket) { perm[0] = 2; perm[1] = 3; perm[2] = 0; perm[3] = 1; }
  if (swap_ij) { int t = perm[0]; perm[0] = perm[1]; perm[1] = t; }
  if (swap_kl) { int t = perm[2]; perm[2] = perm[3]; perm[3] = t; }
  const int d0 = dims[perm[0]], d1 = dims[perm[1]], d2 = dims[perm[2]];
#pragma unroll
  for (int fi = 0; fi < NI; ++fi)
#pragma unroll
    for (int fj = 0; fj < NJ; ++fj)
#pragma unroll
      for (int fk = 0; fk < NK; ++fk)
#pragma unroll
        for (int fl = 0; fl < NL; ++fl) {
          int f[4] = {fi, fj, fk, fl};
          int o = f[perm[0]] + d0 * (f[perm[1]] + d1 * (f[perm[2]] + d2 * f[perm[3]]));
          dst[o] = acc[((fi * NJ + fj) * NK + fk) * NL + fl];
        }
}

template <typename T>
__global__ void __launch_bounds__(256) eri_quartets_kernel(const int* __restrict__ atm, const int* __restrict__ bas,
                                                           const double* __restrict__ env, int nbas,
                                                           const int* __restrict__ quartets,
                                                           const int64_t* __restrict__ offsets, int nq,
                                                           T* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* tab64 = reinterpret_cast<double*>(smem);
  float* tab32 = reinterpret_cast<float*>(smem + sizeof(double) * BOYS_NT * BOYS_NM);
  int* ao_loc = reinterpret_cast<int*>(smem + (sizeof(double) + sizeof(float)) * BOYS_NT * BOYS_NM);
  for (int i = threadIdx.x; i < BOYS_NT * BOYS_NM; i += blockDim.x) {
    tab64[i] = g_boys[i];
    tab32[i] = (float)g_boys[i];
  }
  stage_ao_loc(ao_loc, bas, nbas);
  __syncthreads();
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  int sh[4] = {quartets[4 * q], quartets[4 * q + 1], quartets[4 * q + 2], quartets[4 * q + 3]};
  int l[4];
  for (int t = 0; t < 4; ++t) l[t] = bas[sh[t] * BAS_SLOTS + 1];
  bool swap_ij = l[0] < l[1], swap_kl = l[2] < l[3];
  if (swap_ij) { int t = sh[0]; sh[0] = sh[1]; sh[1] = t; t = l[0]; l[0] = l[1]; l[1] = t; }
  if (swap_kl) { int t = sh[2]; sh[2] = sh[3]; sh[3] = t; t = l[2]; l[2] = l[3]; l[3] = t; }
  bool swap_braket = (l[0] + l[1]) < (l[2] + l[3]);
  if (swap_braket) {
    int t0 = sh[0], t1 = sh[1];
    sh[0] = sh[2]; sh[1] = sh[3]; sh[2] = t0; sh[3] = t1;
    t0 = l[0]; t1 = l[1];
    l[0] = l[2]; l[1] = l[3]; l[2] = t0; l[3] = t1;
  }
  ShellD S[4];
  for (int t = 0; t < 4; ++t) S[t] = load_shell(atm, bas, env, ao_loc, sh[t]);
  const int X = l[0] + l[1], Y = l[2] + l[3];
  T* dst = out + offsets[q];
  // The caller-order permutation is applied relative to the canonical order:
  // canonical = braket-swapped then pair-swapped; invert in quartet_list_class.
  if (X == 0) quartet_list_class<T, 0, 0>(S, swap_ij, swap_kl, swap_braket, tab64, tab32, dst);
  else if (X == 1 && Y == 0) quartet_list_class<T, 1, 0>(S, swap_ij, swap_kl, swap_braket, tab64, tab32, dst);
  else if (X == 1 && Y == 1) quartet_list_class<T, 1, 1>(S, swap_ij, swap_kl, swap_braket, tab64, tab32, dst);
  else if (X == 2 && Y == 0) quartet_list_class<T, 2, 0>(S, swap_ij, swap_kl, swap_braket, tab64, tab32, dst);
  else if (X == 2 && Y == 1) quartet_list_class<T, 2, 1>(S, swap_ij, swap_kl, swap_braket, tab64, tab32, dst);
  else quartet_list_class<T, 2, 2>(S, swap_ij, swap_kl, swap_braket, tab64, tab32, dst);
}

// ------------------------------------------------------------------- int1e ---
// S, T (Obara-Saika) and V (Rys over nuclei atom0..atom0+natm-1) blocks of one
// shell pair computed by one 64-lane wave: lanes take (primitive pair, nucleus)
// combinations, a fixed shuffle tree sums them, lane 0 writes both triangles
// (row stride nao).  hcore: write T + V into Tk instead of T and V.
template <int LA, int LB>
__device__ __forceinline__ void int1e_pair_wave_l(const int* __restrict__ atm, int atom0, int natm, const int* __restrict__ bas,
                                       const double* __restrict__ env, const double* __restrict__ tab, int ia, int ib,
                                       int ao_a, int ao_b, int nao, double* __restrict__ S, double* __restrict__ Tk,
                                       double* __restrict__ V, bool hcore) {
  const int lane = threadIdx.x & 63;
  const int* ba = bas + ia * BAS_SLOTS;
  const int* bb = bas + ib * BAS_SLOTS;
  const double* RA = env + atm[ba[0] * ATM_SLOTS + 1];
  const double* RB = env + atm[bb[0] * ATM_SLOTS + 1];
  constexpr int la = LA, lb = LB;
  const int npa = ba[2], npb = bb[2];
  const double* ea = env + ba[5];
  const double* ca = env + ba[6];
  const double* eb = env + bb[5];
  const double* cb = env + bb[6];
  constexpr int na = la ? 3 : 1, nb = lb ? 3 : 1;
  const double fac = (la ? ANG_P : ANG_S) * (lb ? ANG_P : ANG_S);
  double s_acc[9] = {0}, t_acc[9] = {0}, v_acc[9] = {0};
  double AB[3] = {RA[0] - RB[0], RA[1] - RB[1], RA[2] - RB[2]};
  const double rab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
  const int ncombo = npa * npb * natm;
  for (int cidx = lane; cidx < ncombo; cidx += 64) {
    const int iat = cidx % natm, pp = cidx / natm;
    const int pa = pp / npb, pb = pp % npb;
    const double a = ea[pa], b = eb[pb], p = a + b, rp = 1.0 / p, mu = a * b * rp;
    const double cc = ca[pa] * cb[pb] * fac;
    const double ex = exp(-mu * rab2);
    double PA[3], PB[3], P[3];
    for (int d = 0; d < 3; ++d) {
      P[d] = (a * RA[d] + b * RB[d]) * rp;
      PA[d] = P[d] - RA[d];
      PB[d] = P[d] - RB[d];
    }
    if (iat == 0) {  // overlap and kinetic once per primitive pair
      const double prp = PI * rp;
      const double s00 = prp * sqrt(prp) * ex;
      const double h = 0.5 * rp;
      double s1[3][2][4];
      for (int d = 0; d < 3; ++d) {
        s1[d][0][0] = 1.0;
        s1[d][1][0] = PA[d];
        for (int j = 0; j < 3; ++j)
          for (int i = 0; i <= 1; ++i)
            s1[d][i][j + 1] = PB[d] * s1[d][i][j] + h * ((i ? s1[d][i - 1][j] : 0.0) + (j ? j * s1[d][i][j - 1] : 0.0));
      }
      for (int fa = 0; fa < na; ++fa)
        for (int fb = 0; fb < nb; ++fb) {
          int e1[3], e2[3];
          for (int d = 0; d < 3; ++d) {
            e1[d] = la && fa == d;
            e2[d] = lb && fb == d;
          }
          double tt = 0.0;
          for (int d = 0; d < 3; ++d) {
            const int i = e1[d], j = e2[d];
            const double t1 = b * (2 * j + 1) * s1[d][i][j] - 2.0 * b * b * s1[d][i][j + 2];
            double o = 1.0;
            for (int q = 0; q < 3; ++q)
              if (q != d) o *= s1[q][e1[q]][e2[q]];
            tt += t1 * o;
          }
          s_acc[fa * 3 + fb] += cc * s00 * s1[0][e1[0]][e2[0]] * s1[1][e1[1]][e2[1]] * s1[2][e1[2]][e2[2]];
          t_acc[fa * 3 + fb] += cc * s00 * tt;
        }
    }
    const int c = atom0 + iat;
    const double* Cn = env + atm[c * ATM_SLOTS + 1];
    const double Z = (double)atm[c * ATM_SLOTS];
    double PC[3] = {P[0] - Cn[0], P[1] - Cn[1], P[2] - Cn[2]};
    const double tn = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
    constexpr int nr = (la + lb) / 2 + 1;
    double u[2], w[2];
    if (nr == 1) {
      double f[2];
      boys<double, 1, BOYS_TAYLOR>(tn, tab, f);
      rys_from_moments<1>(f, u, w);
    } else {
      double f[4];
      boys<double, 3, BOYS_TAYLOR>(tn, tab, f);
      rys_from_moments<2>(f, u, w);
    }
    const double vpre = 2.0 * PI * rp * ex * cc;
    for (int r = 0; r < nr; ++r) {
      double J[3][2][2];
      const double hr = 0.5 * (1.0 - u[r]) * rp;
      for (int d = 0; d < 3; ++d) {
        const double pa_ = PA[d] - u[r] * PC[d], pb_ = PB[d] - u[r] * PC[d];
        J[d][0][0] = 1.0;
        J[d][1][0] = pa_;
        J[d][0][1] = pb_;
        J[d][1][1] = pb_ * pa_ + hr;
      }
      for (int fa = 0; fa < na; ++fa)
        for (int fb = 0; fb < nb; ++fb) {
          double v = -Z * w[r] * vpre;
          for (int d = 0; d < 3; ++d) v *= J[d][la && fa == d][lb && fb == d];
          v_acc[fa * 3 + fb] += v;
        }
    }
  }
  // wavefront sums by DPP (totals in lane 63), only of the na nb used entries
#pragma unroll
  for (int k = 0; k < 9; ++k)
    if (k / 3 < na && k % 3 < nb) {  // uniform
      s_acc[k] = wave_sum_last(s_acc[k]);
      t_acc[k] = wave_sum_last(t_acc[k]);
      v_acc[k] = wave_sum_last(v_acc[k]);
    }
  if (lane == 63)
    for (int fa = 0; fa < na; ++fa)
      for (int fb = 0; fb < nb; ++fb) {
        const int64_t r = ao_a + fa, c = ao_b + fb;
        S[r * nao + c] = S[c * nao + r] = s_acc[fa * 3 + fb];
        if (hcore) {
          Tk[r * nao + c] = Tk[c * nao + r] = t_acc[fa * 3 + fb] + v_acc[fa * 3 + fb];
        } else {
          Tk[r * nao + c] = Tk[c * nao + r] = t_acc[fa * 3 + fb];
          V[r * nao + c] = V[c * nao + r] = v_acc[fa * 3 + fb];
        }
      }
}

// Shell classes of the pair as compile-time constants: every loop over AO
// components unrolls and the per-lane accumulators stay in registers.
__device__ __forceinline__ void int1e_pair_wave(const int* __restrict__ atm, int atom0, int natm, const int* __restrict__ bas,
                                       const double* __restrict__ env, const double* __restrict__ tab, int ia, int ib,
                                       int ao_a, int ao_b, int nao, double* __restrict__ S, double* __restrict__ Tk,
                                       double* __restrict__ V, bool hcore) {
  const int la = bas[ia * BAS_SLOTS + 1], lb = bas[ib * BAS_SLOTS + 1];
  if (la == 0 && lb == 0)
    int1e_pair_wave_l<0, 0>(atm, atom0, natm, bas, env, tab, ia, ib, ao_a, ao_b, nao, S, Tk, V, hcore);
  else if (la == 1 && lb == 0)
    int1e_pair_wave_l<1, 0>(atm, atom0, natm, bas, env, tab, ia, ib, ao_a, ao_b, nao, S, Tk, V, hcore);
  else if (la == 0 && lb == 1)
    int1e_pair_wave_l<0, 1>(atm, atom0, natm, bas, env, tab, ia, ib, ao_a, ao_b, nao, S, Tk, V, hcore);
  else
    int1e_pair_wave_l<1, 1>(atm, atom0, natm, bas, env, tab, ia, ib, ao_a, ao_b, nao, S, Tk, V, hcore);
}

__device__ inline int64_t tri_root(int64_t q) {
  int64_t ia = (int64_t)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
  while (tri(ia + 1) <= q) ++ia;
  while (tri(ia) > q) --ia;
  return ia;
}

// One wave per shell pair (i >= j).
__global__ void __launch_bounds__(256) int1e_kernel(const int* __restrict__ atm, int natm,
                                                    const int* __restrict__ bas, int nbas,
                                                    const double* __restrict__ env, int nao, double* __restrict__ S,
                                                    double* __restrict__ Tk, double* __restrict__ V) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* tab = reinterpret_cast<double*>(smem);
  int* ao_loc = reinterpret_cast<int*>(smem + sizeof(double) * BOYS_NT * BOYS_NM);
  for (int i = threadIdx.x; i < BOYS_NT * BOYS_NM; i += blockDim.x) tab[i] = g_boys[i];
  stage_ao_loc(ao_loc, bas, nbas);
  __syncthreads();
  const int64_t q = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (q >= tri((int64_t)nbas)) return;
  const int64_t ia = tri_root(q);
  const int ib = (int)(q - tri(ia));
  int1e_pair_wave(atm, 0, natm, bas, env, tab, (int)ia, ib, ao_loc[ia], ao_loc[ib], nao, S, Tk, V, false);
}

// ------------------------------------------------------------ batched form ---
// Shell pair a of class X from a molecule's class-sorted shell list
// (s shells then p shells): ss = tri over s, ps = (p, s) rectangle, pp = tri over p.
template <int X>
__device__ inline void pair_shells(int64_t a, int ns, int np, const int* __restrict__ sl, int& i, int& j) {
  if (X == 1) {
    i = sl[ns + (int)(a / ns)];
    j = sl[(int)(a % ns)];
  } else {
    int64_t t = (int64_t)((sqrt(8.0 * (double)a + 1.0) - 1.0) * 0.5);
    while (tri(t + 1) <= a) ++t;
    while (tri(t) > a) --t;
    const int base = X == 0 ? 0 : ns;
    i = sl[base + (int)t];
    j = sl[base + (int)(a - tri(t))];
  }
}

__host__ __device__ inline int64_t class_pairs(int X, int64_t ns, int64_t np) {
  return X == 0 ? tri(ns) : (X == 1 ? ns * np : tri(np));
}

// One chunk of <= 256 quartets of class (X, Y) of one molecule per workgroup.
template <typename T, int X, int Y>
__global__ void __launch_bounds__(256) eri_s8_batch_kernel(const BatchMol* __restrict__ mols,
                                                           const int* __restrict__ work,
                                                           const int* __restrict__ atm, const int* __restrict__ bas,
                                                           const double* __restrict__ env,
                                                           const int* __restrict__ shell_ao,
                                                           const int* __restrict__ shl_sorted, T* __restrict__ out,
                                                           int div) {
  constexpr int LI = PairL<X>::a, LJ = PairL<X>::b, LK = PairL<Y>::a, LL = PairL<Y>::b;
  constexpr int NR = (LI + LJ + LK + LL) / 2 + 1;
  using BT = typename BoysType<T, NR>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  BT* tab = reinterpret_cast<BT*>(smem);
  for (int i = threadIdx.x; i < BOYS_NT * BOYS_NM; i += blockDim.x) tab[i] = (BT)g_boys[i];
  __syncthreads();
  const int m = work[2 * blockIdx.x];
  const int64_t q = (int64_t)work[2 * blockIdx.x + 1] + threadIdx.x;
  const BatchMol M = mols[m];
  const int ns = (int)M.ns, np = (int)M.np;
  const int64_t nx = class_pairs(X, ns, np), ny = class_pairs(Y, ns, np);
  const bool tri_launch = (X == Y);
  const int64_t total = tri_launch ? tri(nx) : nx * ny;
  if (q >= total) return;
  int64_t a, b;
  if (tri_launch) {
    a = (int64_t)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
    while (tri(a + 1) <= q) ++a;
    while (tri(a) > q) --a;
    b = q - tri(a);
  } else {
    a = q / ny;
    b = q - a * ny;
  }
  const int* sl = shl_sorted + M.shl0;
  int i, j, k, l;
  pair_shells<X>(a, ns, np, sl, i, j);
  pair_shells<Y>(b, ns, np, sl, k, l);
  // shell_ao holds each shell's first AO within its molecule
  const ShellD A = load_shell(atm, bas, env, shell_ao, i), B = load_shell(atm, bas, env, shell_ao, j);
  const ShellD C = load_shell(atm, bas, env, shell_ao, k), D = load_shell(atm, bas, env, shell_ao, l);
  constexpr int NI = NComp<LI>::v, NJ = NComp<LJ>::v, NK = NComp<LK>::v, NL = NComp<LL>::v;
  T acc[NI * NJ * NK * NL];
  shell_quartet<T, BT, BoysTerms<BT>::v, LI, LJ, LK, LL, true>(A, B, C, D, tab, acc);
  const bool diag_bra = (i == j), diag_ket = (k == l), same_pair = tri_launch && (a == b);
  T* dst = out + M.eri0;
#pragma unroll
  for (int fi = 0; fi < NI; ++fi)
#pragma unroll
    for (int fj = 0; fj < NJ; ++fj)
#pragma unroll
      for (int fk = 0; fk < NK; ++fk)
#pragma unroll
        for (int fl = 0; fl < NL; ++fl) {
          const int64_t p = A.ao0 + fi, qq = B.ao0 + fj, r = C.ao0 + fk, s = D.ao0 + fl;
          if (diag_bra && qq > p) continue;
          if (diag_ket && s > r) continue;
          const int64_t pq = p >= qq ? tri(p) + qq : tri(qq) + p;
          const int64_t rs = r >= s ? tri(r) + s : tri(s) + r;
          if (same_pair && rs > pq) continue;
          const int64_t idx = pq >= rs ? tri(pq) + rs : tri(rs) + pq;
          T v = acc[((fi * NJ + fj) * NK + fk) * NL + fl];
          // div: over the element's multiplicity 2^([p==q] + [r==s] + [pq==rs])
          if (div) v *= T(1.0 / (double)(1 << ((p == qq) + (r == s) + (pq == rs))));
          dst[idx] = v;
        }
}

// One workgroup per molecule, one wave per shell pair: S and H_core = T + V.
__global__ void __launch_bounds__(256) int1e_batch_kernel(const BatchMol* __restrict__ mols,
                                                          const int* __restrict__ atm, const int* __restrict__ bas,
                                                          const double* __restrict__ env,
                                                          const int* __restrict__ shell_ao, double* __restrict__ S,
                                                          double* __restrict__ H) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* tab = reinterpret_cast<double*>(smem);
  for (int i = threadIdx.x; i < BOYS_NT * BOYS_NM; i += blockDim.x) tab[i] = g_boys[i];
  __syncthreads();
  const BatchMol M = mols[blockIdx.x];
  const int nb = (int)M.nbas, b0 = (int)M.bas0, N = (int)M.nao;
  const int nw = blockDim.x >> 6;
  for (int64_t q = threadIdx.x >> 6; q < tri((int64_t)nb); q += nw) {
    const int64_t ia = tri_root(q);
    const int ib = (int)(q - tri(ia));
    const int ga = b0 + (int)ia, gb = b0 + ib;
    int1e_pair_wave(atm, (int)M.atm0, (int)M.natm, bas, env, tab, ga, gb, shell_ao[ga], shell_ao[gb], N,
                    S + M.mat0, H + M.mat0, nullptr, true);
  }
}

// Quartets of pair classes (X, Y): all pairs of X against all of Y, or the
// lower triangle when X == Y.
static int64_t s8_class_quartets(int X, int Y, const int* cnt) {
  return X == Y ? tri((int64_t)cnt[X]) : (int64_t)cnt[X] * cnt[Y];
}

// ------------------------------------------------------------ gradients ---
// The two-electron part of the nuclear gradient (grad_elec,
// pyscf_ipu/nanoDFT/nanoDFT.py:403-433, where int2e_ip1 is contracted with dm0
// into vj and vk): E_2e = 1/2 sum (ij|kl) G_ijkl with
//   G_ijkl = D_ij D_kl - (hyb/4) (D_ik D_jl + D_il D_jk),
// so dE_2e/dR = 1/2 sum over unique shell quartets of their orbit size
// 8 / 2^([i=j] + [k=l] + [ij=kl]) times sum_components G d(ij|kl)/dR.  The
// derivative of a cartesian Gaussian with respect to its centre along d is
// 2a (e + 1_d) - e_d (e - 1_d), so the Rys 2D integrals are built one unit
// higher in n and m than the energy needs; the centres of shells i, j and k
// are differentiated, the centre of l takes minus their sum (translational
// invariance).  FP64 only.

// Horizontal transfer for one axis, j <= 2 and l <= 1: (i j | k l) from the
// moments g(n, m) = (n 0 | m 0), with (x - B)^j = sum_a C(j, a) (x - A)^a AB^(j-a)
// and (x - D)^l likewise with CD.  The indices are compile-time constants
// after unrolling, so the branches fold away.
template <int NG, int MG>
__device__ inline double hrr_axis(const double (&g)[NG][MG], int i, int j, int k, int l, double ab, double cd) {
  auto col = [&](int n) { return l == 0 ? g[n][k] : g[n][k + 1] + cd * g[n][k]; };
  if (j == 0) return col(i);
  if (j == 1) return col(i + 1) + ab * col(i);
  return col(i + 2) + 2.0 * ab * col(i + 1) + ab * ab * col(i);
}

template <int LI, int LJ, int LK, int LL>
__device__ inline void shell_quartet_grad(const ShellD& A, const ShellD& B, const ShellD& C, const ShellD& D,
                                          const double* __restrict__ tab, const double* __restrict__ gam,
                                          double (&gx)[3][3]) {
  constexpr int LT = LI + LJ + LK + LL + 1;
  constexpr int NR = LT / 2 + 1;
  constexpr int NMAX = LI + LJ + 1, MMAX = LK + LL + 1;
  constexpr int NI = NComp<LI>::v, NJ = NComp<LJ>::v, NK = NComp<LK>::v, NL = NComp<LL>::v;
  const double angc = (LI ? ANG_P : ANG_S) * (LJ ? ANG_P : ANG_S) * (LK ? ANG_P : ANG_S) * (LL ? ANG_P : ANG_S);
#pragma unroll
  for (int x = 0; x < 3; ++x)
#pragma unroll
    for (int d = 0; d < 3; ++d) gx[x][d] = 0.0;
  double AB[3], CD[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    AB[d] = A.r[d] - B.r[d];
    CD[d] = C.r[d] - D.r[d];
  }
  const double rab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
  const double rcd2 = CD[0] * CD[0] + CD[1] * CD[1] + CD[2] * CD[2];
#pragma unroll 1
  for (int pa = 0; pa < A.np; ++pa) {
    const double ai = A.e[pa];
#pragma unroll 1
    for (int pb = 0; pb < B.np; ++pb) {
      const double aj = B.e[pb];
      const double aij = ai + aj, r_aij = 1.0 / aij;
      const double kab = exp(-ai * aj * r_aij * rab2) * A.c[pa] * B.c[pb];
      double PA[3], P[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        PA[d] = -aj * r_aij * AB[d];
        P[d] = A.r[d] + PA[d];
      }
#pragma unroll 1
      for (int pc = 0; pc < C.np; ++pc) {
        const double ak = C.e[pc];
#pragma unroll 1
        for (int pd = 0; pd < D.np; ++pd) {
          const double al = D.e[pd];
          const double akl = ak + al, r_akl = 1.0 / akl, alr = al * r_akl;
          const double kcd = exp(-ak * alr * rcd2) * C.c[pc] * D.c[pd];
          double QC[3], PQ[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            QC[d] = -alr * CD[d];
            PQ[d] = P[d] - (C.r[d] + QC[d]);
          }
          const double r_s = 1.0 / (aij + akl);
          const double tt = aij * akl * r_s * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
          const double pref = 34.986836655249725 * r_aij * r_akl * sqrt(r_s) * kab * kcd * angc;
          double f[2 * NR];
          boys<double, 2 * NR - 1, BOYS_TAYLOR>(tt, tab, f);
          double u[NR], w[NR];
          if constexpr (NR == 1) {
            u[0] = f[1] / f[0];
            w[0] = f[0];
          } else {
            rys_from_moments<NR>(f, u, w);
          }
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const double t2 = u[r];
            const double b00 = 0.5 * t2 * r_s;
            const double b10 = 0.5 * r_aij * (1.0 - akl * t2 * r_s);
            const double b01 = 0.5 * r_akl * (1.0 - aij * t2 * r_s);
            // I[d][i][j][k][l]: i <= LI + 1, j <= LJ + 1, k <= LK + 1, l <= LL,
            // at most one of i, j, k above its shell's l (the others unused)
            double I[3][LI + 2][LJ + 2][LK + 2][LL + 1];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              const double c00 = PA[d] - akl * PQ[d] * t2 * r_s;
              const double c0p = QC[d] + aij * PQ[d] * t2 * r_s;
              double g[NMAX + 1][MMAX + 1];
              g[0][0] = 1.0;
#pragma unroll
              for (int n = 0; n < NMAX; ++n) g[n + 1][0] = c00 * g[n][0] + (n > 0 ? n * b10 * g[n - 1][0] : 0.0);
#pragma unroll
              for (int m = 0; m < MMAX; ++m)
#pragma unroll
                for (int n = 0; n <= NMAX; ++n)
                  g[n][m + 1] = c0p * g[n][m] + (m > 0 ? m * b01 * g[n][m - 1] : 0.0) +
                                (n > 0 ? n * b00 * g[n - 1][m] : 0.0);
#pragma unroll
              for (int i = 0; i <= LI + 1; ++i)
#pragma unroll
                for (int j = 0; j <= LJ + 1; ++j)
#pragma unroll
                  for (int k = 0; k <= LK + 1; ++k)
#pragma unroll
                    for (int l = 0; l <= LL; ++l) {
                      if ((i > LI) + (j > LJ) + (k > LK) > 1) continue;
                      I[d][i][j][k][l] = hrr_axis(g, i, j, k, l, AB[d], CD[d]);
                    }
            }
            const double wr = w[r] * pref;
            double sA[3] = {0.0, 0.0, 0.0}, sB[3] = {0.0, 0.0, 0.0}, sC[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int fi = 0; fi < NI; ++fi)
#pragma unroll
              for (int fj = 0; fj < NJ; ++fj)
#pragma unroll
                for (int fk = 0; fk < NK; ++fk)
#pragma unroll
                  for (int fl = 0; fl < NL; ++fl) {
                    const double G = gam[((fi * NJ + fj) * NK + fk) * NL + fl];
                    double base[3];
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                      base[d] = I[d][cexp<LI>(fi, d)][cexp<LJ>(fj, d)][cexp<LK>(fk, d)][cexp<LL>(fl, d)];
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                      const int ei = cexp<LI>(fi, d), ej = cexp<LJ>(fj, d), ek = cexp<LK>(fk, d), el = cexp<LL>(fl, d);
                      const double rest = G * base[(d + 1) % 3] * base[(d + 2) % 3];
                      double dA = 2.0 * ai * I[d][ei + 1][ej][ek][el];
                      if (ei > 0) dA -= ei * I[d][ei - 1][ej][ek][el];
                      double dB = 2.0 * aj * I[d][ei][ej + 1][ek][el];
                      if (ej > 0) dB -= ej * I[d][ei][ej - 1][ek][el];
                      double dC = 2.0 * ak * I[d][ei][ej][ek + 1][el];
                      if (ek > 0) dC -= ek * I[d][ei][ej][ek - 1][el];
                      sA[d] += rest * dA;
                      sB[d] += rest * dB;
                      sC[d] += rest * dC;
                    }
                  }
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              gx[0][d] += wr * sA[d];
              gx[1][d] += wr * sB[d];
              gx[2][d] += wr * sC[d];
            }
          }
        }
      }
    }
  }
}

// Order-independent accumulation of the gradient: each contribution is rounded
// to a multiple of 2^-40 first, so the FP64 sums are exact (below 2^13 in
// magnitude) whatever the order of the atomic adds.
__device__ __forceinline__ double grad_quant(double v) { return rint(v * 0x1p40) * 0x1p-40; }

template <int X, int Y>
__device__ inline void s8_quartet_grad(const int* __restrict__ bas, const double* __restrict__ env,
                                       const int* __restrict__ atm, const int* __restrict__ ao_loc,
                                       const int* __restrict__ pairs, int x0, int nx, int y0, int ny, int64_t q,
                                       const double* tab, const double* __restrict__ dm, int N, double hq,
                                       double* __restrict__ g_lds) {
  constexpr int LI = PairL<X>::a, LJ = PairL<X>::b, LK = PairL<Y>::a, LL = PairL<Y>::b;
  constexpr int NI = NComp<LI>::v, NJ = NComp<LJ>::v, NK = NComp<LK>::v, NL = NComp<LL>::v;
  const bool tri_launch = (X == Y);
  const int64_t total = tri_launch ? tri((int64_t)nx) : (int64_t)nx * ny;
  if (q >= total) return;
  int64_t a, b;
  if (tri_launch) {
    a = (int64_t)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
    while (tri(a + 1) <= q) ++a;
    while (tri(a) > q) --a;
    b = q - tri(a);
  } else {
    a = q / ny;
    b = q - a * ny;
  }
  const int i = pairs[2 * (x0 + a)], j = pairs[2 * (x0 + a) + 1];
  const int k = pairs[2 * (y0 + b)], l = pairs[2 * (y0 + b) + 1];
  const ShellD A = load_shell(atm, bas, env, ao_loc, i), B = load_shell(atm, bas, env, ao_loc, j);
  const ShellD C = load_shell(atm, bas, env, ao_loc, k), D = load_shell(atm, bas, env, ao_loc, l);
  double gam[NI * NJ * NK * NL];
#pragma unroll
  for (int fi = 0; fi < NI; ++fi)
#pragma unroll
    for (int fj = 0; fj < NJ; ++fj)
#pragma unroll
      for (int fk = 0; fk < NK; ++fk)
#pragma unroll
        for (int fl = 0; fl < NL; ++fl) {
          const int p = A.ao0 + fi, qq = B.ao0 + fj, r = C.ao0 + fk, s = D.ao0 + fl;
          gam[((fi * NJ + fj) * NK + fk) * NL + fl] =
              dm[p * N + qq] * dm[r * N + s] - hq * (dm[p * N + r] * dm[qq * N + s] + dm[p * N + s] * dm[qq * N + r]);
        }
  double gx[3][3];
  shell_quartet_grad<LI, LJ, LK, LL>(A, B, C, D, tab, gam, gx);
  const int orbit = (i == j ? 1 : 2) * (k == l ? 1 : 2) * (tri_launch && a == b ? 1 : 2);
  const double f = 0.5 * orbit;
  const int at[4] = {bas[i * BAS_SLOTS], bas[j * BAS_SLOTS], bas[k * BAS_SLOTS], bas[l * BAS_SLOTS]};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double gA = f * gx[0][d], gB = f * gx[1][d], gC = f * gx[2][d];
    atomicAdd(g_lds + at[0] * 3 + d, grad_quant(gA));
    atomicAdd(g_lds + at[1] * 3 + d, grad_quant(gB));
    atomicAdd(g_lds + at[2] * 3 + d, grad_quant(gC));
    atomicAdd(g_lds + at[3] * 3 + d, grad_quant(-(gA + gB + gC)));
  }
}

// One launch per shell-pair class (X, Y), so each class has its own register
// budget: workgroups take 256 consecutive quartets of the class.
template <int X, int Y>
__global__ void __launch_bounds__(256) eri_grad_kernel(const int* __restrict__ atm, int natm,
                                                       const int* __restrict__ bas, const double* __restrict__ env,
                                                       int nbas, const int* __restrict__ pairs, int x0, int nx, int y0,
                                                       int ny, const double* __restrict__ dm, int N, double hq,
                                                       double* __restrict__ grad) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* tab = reinterpret_cast<double*>(smem);
  double* g_lds = tab + BOYS_NT * BOYS_NM;
  int* ao_loc = reinterpret_cast<int*>(g_lds + 3 * natm);
  for (int i = threadIdx.x; i < BOYS_NT * BOYS_NM; i += blockDim.x) tab[i] = g_boys[i];
  for (int i = threadIdx.x; i < 3 * natm; i += blockDim.x) g_lds[i] = 0.0;
  stage_ao_loc(ao_loc, bas, nbas);
  __syncthreads();
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  s8_quartet_grad<X, Y>(bas, env, atm, ao_loc, pairs, x0, nx, y0, ny, q, tab, dm, N, hq, g_lds);
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * natm; i += blockDim.x)
    if (g_lds[i] != 0.0) atomicAdd(grad + i, g_lds[i]);
}

template <typename T>
static int eri_s8_all(const int* atm, const int* bas, const double* env, int nbas, const int* pairs, const int* cnt,
                      T* out, hipStream_t st) {
  static const int order[S8_CLASSES][2] = {{2, 2}, {2, 1}, {2, 0}, {1, 1}, {1, 0}, {0, 0}};
  S8Launch L;
  L.off[0] = 0;
  L.off[1] = cnt[0];
  L.off[2] = cnt[0] + cnt[1];
  for (int x = 0; x < 3; ++x) L.cnt[x] = cnt[x];
  int64_t blocks = 0;
  for (int c = 0; c < S8_CLASSES; ++c) {
    L.b0[c] = (int)blocks;
    blocks += (s8_class_quartets(order[c][0], order[c][1], cnt) + 255) / 256;
    PSCF_REQUIRE(blocks < (1ll << 31), "pscf_eri_s8: too many quartets");
  }
  L.b0[S8_CLASSES] = (int)blocks;
  if (blocks == 0) return PSCF_OK;
  const size_t shm = (sizeof(double) + sizeof(float)) * BOYS_NT * BOYS_NM + sizeof(int) * (nbas + 1);
  PSCF_SET_LDS(eri_s8_kernel<T>, shm);
  eri_s8_kernel<T><<<(unsigned)blocks, 256, shm, st>>>(atm, bas, env, nbas, pairs, L, out);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

// One-electron part of the nuclear gradient (the h1aos and s1 terms of
// grad_elec, pyscf_ipu/nanoDFT/nanoDFT.py:422-433, with hcore_deriv
// :463-470).  Over ordered shell pairs (i, j), with d_A the derivative with
// respect to the centre A of shell i (2a (e + 1_d) - e_d (e - 1_d) on its
// cartesian exponents):
//   grad[atom(i)] += 2 sum D_ij d_A (T + V)_ij - 2 sum W_ij d_A S_ij,
//   grad[C]       -= 2 sum D_ij d_A V^C_ij       for every nucleus C,
// where V = sum_C V^C, V^C = -Z_C <i|1/|r - C||j>: its derivative with respect
// to C is -(d_A + d_B) V^C, and summing over ordered pairs folds d_B into d_A
// (D and the integrals are symmetric).  W is the energy-weighted density.
template <int LA, int LB>
__device__ inline void int1e_grad_pair(const ShellD& A, const ShellD& B, const int* __restrict__ atm, int natm,
                                       const double* __restrict__ env, const double* __restrict__ tab,
                                       const double* __restrict__ dm, const double* __restrict__ dme, int N,
                                       double (&gA)[3], double* __restrict__ g_lds) {
  constexpr int NA = NComp<LA>::v, NB = NComp<LB>::v;
  constexpr int NR = (LA + 1 + LB) / 2 + 1;
  const double angc = (LA ? ANG_P : ANG_S) * (LB ? ANG_P : ANG_S);
  double Dab[NA][NB], Wab[NA][NB];
#pragma unroll
  for (int fa = 0; fa < NA; ++fa)
#pragma unroll
    for (int fb = 0; fb < NB; ++fb) {
      Dab[fa][fb] = dm[(A.ao0 + fa) * N + B.ao0 + fb];
      Wab[fa][fb] = dme[(A.ao0 + fa) * N + B.ao0 + fb];
    }
  double AB[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) AB[d] = A.r[d] - B.r[d];
  const double rab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
#pragma unroll
  for (int d = 0; d < 3; ++d) gA[d] = 0.0;
#pragma unroll 1
  for (int pa = 0; pa < A.np; ++pa) {
#pragma unroll 1
    for (int pb = 0; pb < B.np; ++pb) {
      const double a = A.e[pa], b = B.e[pb], p = a + b, rp = 1.0 / p, mu = a * b * rp;
      const double cc = A.c[pa] * B.c[pb] * angc;
      const double eab = exp(-mu * rab2);
      double P[3], PA[3], PB[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        P[d] = (a * A.r[d] + b * B.r[d]) * rp;
        PA[d] = P[d] - A.r[d];
        PB[d] = P[d] - B.r[d];
      }
      // overlap moments S[d][i][j], i <= LA + 1, j <= LB + 2 (Obara-Saika)
      double S[3][LA + 2][LB + 3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        S[d][0][0] = 1.0;
#pragma unroll
        for (int i = 0; i <= LA; ++i) S[d][i + 1][0] = PA[d] * S[d][i][0] + (i > 0 ? i * 0.5 * rp * S[d][i - 1][0] : 0.0);
#pragma unroll
        for (int j = 0; j <= LB + 1; ++j)
#pragma unroll
          for (int i = 0; i <= LA + 1; ++i)
            S[d][i][j + 1] = PB[d] * S[d][i][j] +
                             0.5 * rp * ((i > 0 ? i * S[d][i - 1][j] : 0.0) + (j > 0 ? j * S[d][i][j - 1] : 0.0));
      }
      // kinetic factor of one axis: b (2j + 1) S(i, j) - 2 b^2 S(i, j + 2) - j (j - 1) / 2 S(i, j - 2)
      auto kin = [&](int d, int i, int j) {
        return b * (2 * j + 1) * S[d][i][j] - 2.0 * b * b * S[d][i][j + 2] - (j >= 2 ? 0.5 * j * (j - 1) * S[d][i][j - 2] : 0.0);
      };
      const double s00 = cc * PI * rp * sqrt(PI * rp) * eab;
      double dS[3] = {0.0, 0.0, 0.0}, dT[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int fa = 0; fa < NA; ++fa)
#pragma unroll
        for (int fb = 0; fb < NB; ++fb) {
          int ea[3], eb[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            ea[d] = cexp<LA>(fa, d);
            eb[d] = cexp<LB>(fb, d);
          }
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
            const double sd = 2.0 * a * S[d][ea[d] + 1][eb[d]] - (ea[d] > 0 ? ea[d] * S[d][ea[d] - 1][eb[d]] : 0.0);
            const double td = 2.0 * a * kin(d, ea[d] + 1, eb[d]) - (ea[d] > 0 ? ea[d] * kin(d, ea[d] - 1, eb[d]) : 0.0);
            const double s1 = S[d1][ea[d1]][eb[d1]], s2 = S[d2][ea[d2]][eb[d2]];
            dS[d] += Wab[fa][fb] * sd * s1 * s2;
            dT[d] += Dab[fa][fb] * (td * s1 * s2 + sd * (kin(d1, ea[d1], eb[d1]) * s2 + s1 * kin(d2, ea[d2], eb[d2])));
          }
        }
#pragma unroll
      for (int d = 0; d < 3; ++d) gA[d] += 2.0 * s00 * (dT[d] - dS[d]);
      // nuclear attraction, Rys quadrature over each nucleus
      const double vpre = cc * 2.0 * PI * rp * eab;
      for (int c = 0; c < natm; ++c) {
        const double* Cn = env + atm[c * ATM_SLOTS + 1];
        const double Z = (double)atm[c * ATM_SLOTS];
        double PC[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) PC[d] = P[d] - Cn[d];
        const double tn = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
        double f[2 * NR], u[NR], w[NR];
        boys<double, 2 * NR - 1, BOYS_TAYLOR>(tn, tab, f);
        rys_from_moments<NR>(f, u, w);
        double dV[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const double h = 0.5 * (1.0 - u[r]) * rp;
          double J[3][LA + 2][LB + 1];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double pa_ = PA[d] - u[r] * PC[d], pb_ = PB[d] - u[r] * PC[d];
            J[d][0][0] = 1.0;
#pragma unroll
            for (int i = 0; i <= LA; ++i) J[d][i + 1][0] = pa_ * J[d][i][0] + (i > 0 ? i * h * J[d][i - 1][0] : 0.0);
#pragma unroll
            for (int j = 0; j < LB; ++j)
#pragma unroll
              for (int i = 0; i <= LA + 1; ++i)
                J[d][i][j + 1] = pb_ * J[d][i][j] + h * ((i > 0 ? i * J[d][i - 1][j] : 0.0) + (j > 0 ? j * J[d][i][j - 1] : 0.0));
          }
#pragma unroll
          for (int fa = 0; fa < NA; ++fa)
#pragma unroll
            for (int fb = 0; fb < NB; ++fb)
#pragma unroll
              for (int d = 0; d < 3; ++d) {
                const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
                const int ed = cexp<LA>(fa, d), fd = cexp<LB>(fb, d);
                const double jd = 2.0 * a * J[d][ed + 1][fd] - (ed > 0 ? ed * J[d][ed - 1][fd] : 0.0);
                dV[d] += w[r] * Dab[fa][fb] * jd * J[d1][cexp<LA>(fa, d1)][cexp<LB>(fb, d1)] *
                         J[d2][cexp<LA>(fa, d2)][cexp<LB>(fb, d2)];
              }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double t = 2.0 * (-Z) * vpre * dV[d];  // 2 D d_A V^C
          gA[d] += t;
          atomicAdd(g_lds + c * 3 + d, grad_quant(-t));
        }
      }
    }
  }
}

// Lane per ordered shell pair (i, j), i = pair / nbas.
__global__ void __launch_bounds__(256) int1e_grad_kernel(const int* __restrict__ atm, int natm,
                                                         const int* __restrict__ bas, const double* __restrict__ env,
                                                         int nbas, const double* __restrict__ dm,
                                                         const double* __restrict__ dme, int N,
                                                         double* __restrict__ grad) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* tab = reinterpret_cast<double*>(smem);
  double* g_lds = tab + BOYS_NT * BOYS_NM;
  int* ao_loc = reinterpret_cast<int*>(g_lds + 3 * natm);
  for (int i = threadIdx.x; i < BOYS_NT * BOYS_NM; i += blockDim.x) tab[i] = g_boys[i];
  for (int i = threadIdx.x; i < 3 * natm; i += blockDim.x) g_lds[i] = 0.0;
  stage_ao_loc(ao_loc, bas, nbas);
  __syncthreads();
  const int64_t pr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pr < (int64_t)nbas * nbas) {
    const int i = (int)(pr / nbas), j = (int)(pr - (int64_t)i * nbas);
    const ShellD A = load_shell(atm, bas, env, ao_loc, i), B = load_shell(atm, bas, env, ao_loc, j);
    const int la = bas[i * BAS_SLOTS + 1], lb = bas[j * BAS_SLOTS + 1];
    double gA[3];
    if (la == 0 && lb == 0) int1e_grad_pair<0, 0>(A, B, atm, natm, env, tab, dm, dme, N, gA, g_lds);
    else if (la == 0) int1e_grad_pair<0, 1>(A, B, atm, natm, env, tab, dm, dme, N, gA, g_lds);
    else if (lb == 0) int1e_grad_pair<1, 0>(A, B, atm, natm, env, tab, dm, dme, N, gA, g_lds);
    else int1e_grad_pair<1, 1>(A, B, atm, natm, env, tab, dm, dme, N, gA, g_lds);
    const int at = bas[i * BAS_SLOTS];
#pragma unroll
    for (int d = 0; d < 3; ++d) atomicAdd(g_lds + at * 3 + d, grad_quant(gA[d]));
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * natm; i += blockDim.x)
    if (g_lds[i] != 0.0) atomicAdd(grad + i, g_lds[i]);
}

// The six shell-pair classes of the class-sorted pair list of pscf_eri_s8,
// one launch each.
template <int X, int Y>
static int eri_grad_class(const int* atm, int natm, const int* bas, const double* env, int nbas, const int* pairs,
                          const int* cnt, const double* dm, int N, double hyb, double* grad, hipStream_t st) {
  const int off[3] = {0, cnt[0], cnt[0] + cnt[1]};
  const int64_t nq = s8_class_quartets(X, Y, cnt);
  if (nq == 0) return PSCF_OK;
  const int64_t blocks = (nq + 255) / 256;
  PSCF_REQUIRE(blocks < (1ll << 31), "pscf_grad_eri: too many quartets");
  const size_t shm = sizeof(double) * (BOYS_NT * BOYS_NM + 3 * (size_t)natm) + sizeof(int) * (nbas + 1);
  PSCF_SET_LDS((eri_grad_kernel<X, Y>), shm);
  eri_grad_kernel<X, Y><<<(unsigned)blocks, 256, shm, st>>>(atm, natm, bas, env, nbas, pairs, off[X], cnt[X], off[Y],
                                                            cnt[Y], dm, N, hyb / 4.0, grad);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

static int eri_grad_all(const int* atm, int natm, const int* bas, const double* env, int nbas, const int* pairs,
                        const int* cnt, const double* dm, int N, double hyb, double* grad, hipStream_t st) {
  int rc = eri_grad_class<2, 2>(atm, natm, bas, env, nbas, pairs, cnt, dm, N, hyb, grad, st);
  if (rc == PSCF_OK) rc = eri_grad_class<2, 1>(atm, natm, bas, env, nbas, pairs, cnt, dm, N, hyb, grad, st);
  if (rc == PSCF_OK) rc = eri_grad_class<2, 0>(atm, natm, bas, env, nbas, pairs, cnt, dm, N, hyb, grad, st);
  if (rc == PSCF_OK) rc = eri_grad_class<1, 1>(atm, natm, bas, env, nbas, pairs, cnt, dm, N, hyb, grad, st);
  if (rc == PSCF_OK) rc = eri_grad_class<1, 0>(atm, natm, bas, env, nbas, pairs, cnt, dm, N, hyb, grad, st);
  if (rc == PSCF_OK) rc = eri_grad_class<0, 0>(atm, natm, bas, env, nbas, pairs, cnt, dm, N, hyb, grad, st);
  return rc;
}

}  // namespace pscf

using namespace pscf;

extern "C" int pscf_eri_s8(int dtype, const int* d_atm, int natm, const int* d_bas, int nbas, const double* d_env,
                           const int* d_pairs, const int* h_class_counts, int nao, void* d_out, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_eri_s8: bad dtype %d", dtype);
  PSCF_REQUIRE(natm > 0 && nbas > 0 && nao > 0, "pscf_eri_s8: empty molecule");
  PSCF_REQUIRE(d_atm && d_bas && d_env && d_pairs && h_class_counts && d_out, "pscf_eri_s8: null pointer");
  PSCF_REQUIRE(nbas <= 4096, "pscf_eri_s8: nbas %d > 4096", nbas);
  int64_t npair = (int64_t)h_class_counts[0] + h_class_counts[1] + h_class_counts[2];
  PSCF_REQUIRE(h_class_counts[0] >= 0 && h_class_counts[1] >= 0 && h_class_counts[2] >= 0 &&
                   npair == tri((int64_t)nbas),
               "pscf_eri_s8: class counts must cover all nbas(nbas+1)/2 shell pairs");
  int rc = ensure_boys_table();
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64) return eri_s8_all<double>(d_atm, d_bas, d_env, nbas, d_pairs, h_class_counts, (double*)d_out, st);
  return eri_s8_all<float>(d_atm, d_bas, d_env, nbas, d_pairs, h_class_counts, (float*)d_out, st);
}

extern "C" int pscf_grad_eri(const int* d_atm, int natm, const int* d_bas, int nbas, const double* d_env,
                             const int* d_pairs, const int* h_class_counts, int nao, const double* d_dm, double hyb,
                             double* d_grad, void* stream) {
  PSCF_REQUIRE(natm > 0 && natm <= 4096 && nbas > 0 && nbas <= 4096 && nao > 0, "pscf_grad_eri: bad molecule");
  PSCF_REQUIRE(d_atm && d_bas && d_env && d_pairs && h_class_counts && d_dm && d_grad, "pscf_grad_eri: null pointer");
  int64_t npair = (int64_t)h_class_counts[0] + h_class_counts[1] + h_class_counts[2];
  PSCF_REQUIRE(h_class_counts[0] >= 0 && h_class_counts[1] >= 0 && h_class_counts[2] >= 0 &&
                   npair == tri((int64_t)nbas),
               "pscf_grad_eri: class counts must cover all nbas(nbas+1)/2 shell pairs");
  int rc = ensure_boys_table();
  if (rc) return rc;
  return eri_grad_all(d_atm, natm, d_bas, d_env, nbas, d_pairs, h_class_counts, d_dm, nao, hyb, d_grad,
                      (hipStream_t)stream);
}

extern "C" int pscf_grad_int1e(const int* d_atm, int natm, const int* d_bas, int nbas, const double* d_env, int nao,
                               const double* d_dm, const double* d_dme, double* d_grad, void* stream) {
  PSCF_REQUIRE(natm > 0 && natm <= 4096 && nbas > 0 && nbas <= 4096 && nao > 0, "pscf_grad_int1e: bad molecule");
  PSCF_REQUIRE(d_atm && d_bas && d_env && d_dm && d_dme && d_grad, "pscf_grad_int1e: null pointer");
  int rc = ensure_boys_table();
  if (rc) return rc;
  const int64_t npairs = (int64_t)nbas * nbas;
  const size_t shm = sizeof(double) * (BOYS_NT * BOYS_NM + 3 * (size_t)natm) + sizeof(int) * (nbas + 1);
  PSCF_SET_LDS(int1e_grad_kernel, shm);
  int1e_grad_kernel<<<(unsigned)((npairs + 255) / 256), 256, shm, (hipStream_t)stream>>>(d_atm, natm, d_bas, d_env,
                                                                                        nbas, d_dm, d_dme, nao,
                                                                                        d_grad);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_eri_quartets(int dtype, const int* d_atm, int natm, const int* d_bas, int nbas,
                                 const double* d_env, const int* d_quartets, const int64_t* d_offsets, int nq,
                                 void* d_out, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_eri_quartets: bad dtype %d", dtype);
  PSCF_REQUIRE(natm > 0 && nbas > 0 && nbas <= 4096, "pscf_eri_quartets: bad molecule");
  PSCF_REQUIRE(nq >= 0, "pscf_eri_quartets: nq < 0");
  if (nq == 0) return PSCF_OK;
  PSCF_REQUIRE(d_atm && d_bas && d_env && d_quartets && d_offsets && d_out, "pscf_eri_quartets: null pointer");
  int rc = ensure_boys_table();
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const size_t shm = (sizeof(double) + sizeof(float)) * BOYS_NT * BOYS_NM + sizeof(int) * (nbas + 1);
  const int blocks = (nq + 255) / 256;
  PSCF_SET_LDS(eri_quartets_kernel<double>, shm);
  PSCF_SET_LDS(eri_quartets_kernel<float>, shm);
  if (dtype == PSCF_F64)
    eri_quartets_kernel<double><<<blocks, 256, shm, st>>>(d_atm, d_bas, d_env, nbas, d_quartets, d_offsets, nq,
                                                          (double*)d_out);
  else
    eri_quartets_kernel<float><<<blocks, 256, shm, st>>>(d_atm, d_bas, d_env, nbas, d_quartets, d_offsets, nq,
                                                         (float*)d_out);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_int1e(const int* d_atm, int natm, const int* d_bas, int nbas, const double* d_env, int nao,
                          double* d_S, double* d_T, double* d_V, void* stream) {
  PSCF_REQUIRE(natm > 0 && nbas > 0 && nbas <= 4096 && nao > 0, "pscf_int1e: bad molecule");
  PSCF_REQUIRE(d_atm && d_bas && d_env && d_S && d_T && d_V, "pscf_int1e: null pointer");
  int rc = ensure_boys_table();
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int64_t npair = tri((int64_t)nbas);
  const size_t shm = sizeof(double) * BOYS_NT * BOYS_NM + sizeof(int) * (nbas + 1);
  int1e_kernel<<<(unsigned)((npair + 3) / 4), 256, shm, st>>>(d_atm, natm, d_bas, nbas, d_env, nao, d_S, d_T, d_V);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

template <typename T, int X, int Y>
static int launch_batch_s8(const BatchMol* mols, const int* work, int nwork, const int* atm, const int* bas,
                           const double* env, const int* shell_ao, const int* shl, T* out, int div, hipStream_t st) {
  if (nwork == 0) return PSCF_OK;
  constexpr int LI = PairL<X>::a, LJ = PairL<X>::b, LK = PairL<Y>::a, LL = PairL<Y>::b;
  constexpr int NR = (LI + LJ + LK + LL) / 2 + 1;
  using BT = typename BoysType<T, NR>::type;
  const size_t shm = sizeof(BT) * BOYS_NT * BOYS_NM;
  eri_s8_batch_kernel<T, X, Y><<<nwork, 256, shm, st>>>(mols, work, atm, bas, env, shell_ao, shl, out, div);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

template <typename T>
static int batch_s8_all(const BatchMol* mols, const int* const* work, const int* nwork, const int* atm, const int* bas,
                        const double* env, const int* shell_ao, const int* shl, T* out, int div, hipStream_t st) {
  int rc;
  if ((rc = launch_batch_s8<T, 0, 0>(mols, work[0], nwork[0], atm, bas, env, shell_ao, shl, out, div, st))) return rc;
  if ((rc = launch_batch_s8<T, 1, 0>(mols, work[1], nwork[1], atm, bas, env, shell_ao, shl, out, div, st))) return rc;
  if ((rc = launch_batch_s8<T, 1, 1>(mols, work[2], nwork[2], atm, bas, env, shell_ao, shl, out, div, st))) return rc;
  if ((rc = launch_batch_s8<T, 2, 0>(mols, work[3], nwork[3], atm, bas, env, shell_ao, shl, out, div, st))) return rc;
  if ((rc = launch_batch_s8<T, 2, 1>(mols, work[4], nwork[4], atm, bas, env, shell_ao, shl, out, div, st))) return rc;
  if ((rc = launch_batch_s8<T, 2, 2>(mols, work[5], nwork[5], atm, bas, env, shell_ao, shl, out, div, st))) return rc;
  return PSCF_OK;
}

static int batch_eri_s8_any(int dtype, const void* d_mols, int nmol, const int* const* h_work, const int* h_nwork,
                            const int* d_atm, const int* d_bas, const double* d_env, const int* d_shell_ao,
                            const int* d_shl_sorted, void* d_out, int div, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_batch_eri_s8: bad dtype");
  PSCF_REQUIRE(nmol > 0 && d_mols && h_work && h_nwork && d_atm && d_bas && d_env && d_shell_ao && d_shl_sorted &&
                   d_out,
               "pscf_batch_eri_s8: bad arguments");
  for (int c = 0; c < 6; ++c) PSCF_REQUIRE(h_nwork[c] >= 0 && (h_nwork[c] == 0 || h_work[c]), "pscf_batch_eri_s8: bad work list");
  int rc = ensure_boys_table();
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64)
    return batch_s8_all<double>((const BatchMol*)d_mols, h_work, h_nwork, d_atm, d_bas, d_env, d_shell_ao,
                                d_shl_sorted, (double*)d_out, div, st);
  return batch_s8_all<float>((const BatchMol*)d_mols, h_work, h_nwork, d_atm, d_bas, d_env, d_shell_ao, d_shl_sorted,
                             (float*)d_out, div, st);
}

extern "C" int pscf_batch_eri_s8(int dtype, const void* d_mols, int nmol, const int* const* h_work,
                                 const int* h_nwork, const int* d_atm, const int* d_bas, const double* d_env,
                                 const int* d_shell_ao, const int* d_shl_sorted, void* d_out, void* stream) {
  return batch_eri_s8_any(dtype, d_mols, nmol, h_work, h_nwork, d_atm, d_bas, d_env, d_shell_ao, d_shl_sorted, d_out,
                          0, stream);
}

extern "C" int pscf_batch_eri_s8_div(int dtype, const void* d_mols, int nmol, const int* const* h_work,
                                     const int* h_nwork, const int* d_atm, const int* d_bas, const double* d_env,
                                     const int* d_shell_ao, const int* d_shl_sorted, void* d_out, void* stream) {
  return batch_eri_s8_any(dtype, d_mols, nmol, h_work, h_nwork, d_atm, d_bas, d_env, d_shell_ao, d_shl_sorted, d_out,
                          1, stream);
}

// Work list of (owner, first index) pairs covering range(count_i) of every
// owner i in steps of `size`: entry t belongs to the owner i with first[i] <=
// t < first[i + 1], first being the exclusive prefix sum of the owners' chunk
// counts ceil(count_i / size) (first[n] = total).
__global__ void __launch_bounds__(256) chunks_kernel(const int64_t* __restrict__ first, int n, int size,
                                                     int* __restrict__ work, int64_t total) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = n - 1;  // largest i with first[i] <= t
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (first[mid] <= t)
        lo = mid;
      else
        hi = mid - 1;
    }
    work[2 * t] = lo;
    work[2 * t + 1] = (int)((t - first[lo]) * size);
  }
}

extern "C" int pscf_batch_chunks(const int64_t* d_first, int n, int size, int* d_work, int64_t total,
                                 void* stream) {
  PSCF_REQUIRE(n > 0 && size > 0 && total >= 0 && d_first && (total == 0 || d_work), "pscf_batch_chunks: bad arguments");
  if (total == 0) return PSCF_OK;
  const int64_t blocks = (total + 255) / 256;
  chunks_kernel<<<(unsigned)(blocks < 8192 ? blocks : 8192), 256, 0, (hipStream_t)stream>>>(d_first, n, size, d_work,
                                                                                             total);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

// Nuclear repulsion of every molecule, one thread per molecule, atom pairs
// i > j in order (nanoDFT.py: mol.energy_nuc()).
__global__ void __launch_bounds__(256) enuc_kernel(const BatchMol* __restrict__ mols, int nmol,
                                                   const int* __restrict__ atm, const double* __restrict__ env,
                                                   double* __restrict__ out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmol) return;
  const int a0 = (int)mols[m].atm0, na = (int)mols[m].natm;
  double e = 0.0;
  for (int i = 1; i < na; ++i) {
    const double* Ri = env + atm[(a0 + i) * ATM_SLOTS + 1];
    const double Zi = (double)atm[(a0 + i) * ATM_SLOTS];
    for (int j = 0; j < i; ++j) {
      const double* Rj = env + atm[(a0 + j) * ATM_SLOTS + 1];
      const double dx = Ri[0] - Rj[0], dy = Ri[1] - Rj[1], dz = Ri[2] - Rj[2];
      e += Zi * (double)atm[(a0 + j) * ATM_SLOTS] / sqrt(dx * dx + dy * dy + dz * dz);
    }
  }
  out[m] = e;
}

extern "C" int pscf_batch_enuc(const void* d_mols, int nmol, const int* d_atm, const double* d_env, double* d_enuc,
                               void* stream) {
  PSCF_REQUIRE(nmol > 0 && d_mols && d_atm && d_env && d_enuc, "pscf_batch_enuc: bad arguments");
  enuc_kernel<<<(nmol + 255) / 256, 256, 0, (hipStream_t)stream>>>((const BatchMol*)d_mols, nmol, d_atm, d_env,
                                                                    d_enuc);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_batch_int1e(const void* d_mols, int nmol, const int* d_atm, const int* d_bas, const double* d_env,
                                const int* d_shell_ao, double* d_S, double* d_Hcore, void* stream) {
  PSCF_REQUIRE(nmol > 0 && d_mols && d_atm && d_bas && d_env && d_shell_ao && d_S && d_Hcore,
               "pscf_batch_int1e: bad arguments");
  int rc = ensure_boys_table();
  if (rc) return rc;
  const size_t shm = sizeof(double) * BOYS_NT * BOYS_NM;
  int1e_batch_kernel<<<nmol, 256, shm, (hipStream_t)stream>>>((const BatchMol*)d_mols, d_atm, d_bas, d_env,
                                                              d_shell_ao, d_S, d_Hcore);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}
