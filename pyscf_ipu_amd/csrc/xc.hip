// Grid exchange-correlation path on gfx950: AO values/gradients, Becke cell
// weights, B3LYP pointwise, and the fused rho -> B3LYP -> V_xc kernel.
//
// Reference: exchange_correlation (pyscf_ipu/nanoDFT/nanoDFT.py:105-123):
//   X = phi0 D; rho_c = sum_mu X phi_c (rho, grad rho / 2); b3lyp;
//   E_xc = sum rho w e_xc; V = phi0^T sum_c phi_c v_c w with
//   v = [v_rho/2, 4 v_sigma grad rho / 2]; V_xc = V + V^T.
// Fused form: each workgroup walks 64-point tiles of the grid, keeping D and
// the current tile of phi0 in LDS, forming X, rho, the pointwise potential
// and Y = sum_c phi_c v_c w in LDS, and accumulating its share of
// V = phi0^T Y in registers (4x4 register blocks); one partial V per
// workgroup is reduced in a fixed order by a second kernel.
#include <type_traits>

#include "b3lyp.h"
#include "batch.h"
#include "common.h"
#include "reduce.h"
#include "wave.h"

namespace pscf {

static int xc_max_lds() { return device_max_lds(); }

// ---------------------------------------------------------------- AO eval ---
template <typename T>
__global__ void __launch_bounds__(256) ao_eval_kernel(const int* __restrict__ atm, const int* __restrict__ bas,
                                                      int nbas, const double* __restrict__ env,
                                                      const double* __restrict__ coords, int g, int nao,
                                                      T* __restrict__ ao) {
  const int pt = blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= g) return;
  const double x = coords[3 * pt], y = coords[3 * pt + 1], z = coords[3 * pt + 2];
  const int64_t cs = (int64_t)nao * g;
  int mu = 0;
  for (int ib = 0; ib < nbas; ++ib) {
    const int* b = bas + ib * BAS_SLOTS;
    const double* R = env + atm[b[0] * ATM_SLOTS + 1];
    const int l = b[1], np = b[2];
    const double* e = env + b[5];
    const double* c = env + b[6];
    const double dx = x - R[0], dy = y - R[1], dz = z - R[2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    double rad = 0.0, drad = 0.0;
    for (int k = 0; k < np; ++k) {
      const double ek = c[k] * exp(-e[k] * r2);
      rad += ek;
      drad += e[k] * ek;
    }
    drad *= -2.0;
    if (l == 0) {
      rad *= ANG_S;
      drad *= ANG_S;
      ao[(int64_t)mu * g + pt] = (T)rad;
      ao[cs + (int64_t)mu * g + pt] = (T)(drad * dx);
      ao[2 * cs + (int64_t)mu * g + pt] = (T)(drad * dy);
      ao[3 * cs + (int64_t)mu * g + pt] = (T)(drad * dz);
      mu += 1;
    } else {
      rad *= ANG_P;
      drad *= ANG_P;
      const double d[3] = {dx, dy, dz};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int64_t o = (int64_t)(mu + a) * g + pt;
        ao[o] = (T)(d[a] * rad);
#pragma unroll
        for (int q = 0; q < 3; ++q) ao[(1 + q) * cs + o] = (T)((a == q ? rad : 0.0) + d[a] * d[q] * drad);
      }
      mu += 3;
    }
  }
}

// ------------------------------------------------------------ Becke cells ---
// w = vol * P_owner / sum_i P_i with Becke's three-fold smoothed step and
// Treutler's Bragg-radius adjustment a_ij (pyscf_ipu/pyscf_utils/build_grid.py:34-100).
__global__ void __launch_bounds__(256) becke_kernel(const double* __restrict__ coords, const int* __restrict__ owner,
                                                    const double* __restrict__ vol, int g,
                                                    const double* __restrict__ atom_xyz, const double* __restrict__ adj,
                                                    int natm, double* __restrict__ w) {
  extern __shared__ double sh[];
  double* axyz = sh;             // natm * 3
  double* rinv = sh + 3 * natm;  // natm * natm (1 / R_ij)
  const double* a = adj;         // natm * natm, read through the cache
  for (int i = threadIdx.x; i < 3 * natm; i += blockDim.x) axyz[i] = atom_xyz[i];
  for (int i = threadIdx.x; i < natm * natm; i += blockDim.x) {
    const int p = i / natm, q = i % natm;
    const double dx = atom_xyz[3 * p] - atom_xyz[3 * q], dy = atom_xyz[3 * p + 1] - atom_xyz[3 * q + 1],
                 dz = atom_xyz[3 * p + 2] - atom_xyz[3 * q + 2];
    rinv[i] = p == q ? 0.0 : 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
  }
  __syncthreads();
  const int pt = blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= g) return;
  const double x = coords[3 * pt], y = coords[3 * pt + 1], z = coords[3 * pt + 2];
  const int own = owner[pt];
  double psum = 0.0, pown = 0.0;
  for (int i = 0; i < natm; ++i) {
    const double dxi = x - axyz[3 * i], dyi = y - axyz[3 * i + 1], dzi = z - axyz[3 * i + 2];
    const double ri = sqrt(dxi * dxi + dyi * dyi + dzi * dzi);
    double pi = 1.0;
    for (int j = 0; j < natm; ++j) {
      if (j == i) continue;
      const double dxj = x - axyz[3 * j], dyj = y - axyz[3 * j + 1], dzj = z - axyz[3 * j + 2];
      const double rj = sqrt(dxj * dxj + dyj * dyj + dzj * dzj);
      // mu_ij with the adjustment of the ordered pair (i, j)
      double gg = (ri - rj) * rinv[i * natm + j];
      gg = gg + a[i * natm + j] * (1.0 - gg * gg);
      gg = (3.0 - gg * gg) * gg * 0.5;
      gg = (3.0 - gg * gg) * gg * 0.5;
      gg = (3.0 - gg * gg) * gg * 0.5;
      pi *= 0.5 * (1.0 - gg);
    }
    psum += pi;
    if (i == own) pown = pi;
  }
  w[pt] = vol[pt] * pown / psum;
}

// --------------------------------------------------------------- B3LYP -----
template <typename T>
__global__ void b3lyp_kernel(const T* __restrict__ rho, int g, T eps, T* __restrict__ exc, T* __restrict__ vrho,
                             T* __restrict__ vnorm) {
  const int pt = blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= g) return;
  T e, v, vn;
  b3lyp_point<T>(rho[pt], rho[g + pt], rho[2 * g + pt], rho[3 * g + pt], eps, e, v, vn);
  exc[pt] = e;
  vrho[pt] = v;
  vnorm[pt] = vn;
}

// ------------------------------------------------------------- fused XC ----
constexpr int XC_TILE = 64;
constexpr int XC_THREADS = 256;
constexpr int XC_MAX_BLOCKS = 256;

template <typename T>
struct XCArgs {
  const T* ao;
  int64_t sc, smu, sp;
  int g, N;
  const T* dm;
  const T* w;
  T eps;
  T* part_v;        // [blocks][N*N]
  double* part_e;   // [blocks]
  T* vout;          // FP64: V_xc (zeroed) and E_xc take each workgroup's part by
  double* eout;     //   exact atomics instead of the partials (nullptr: partials)
};

// LDS layout: D[N][N] | phi0[N][TILE+1] | XY[N][TILE+1] | coef[4][TILE] | red[8]
// ---- MFMA path ------------------------------------------------------------------
// Per tile of XT points (64 in FP32, 32 in FP64), with the tile's AO values and
// gradients staged in LDS as Phi[c][p][nu] (row stride XS = NP + 4, NP = N
// rounded up to 16):
//   GEMM 1 (v_mfma_{f32,f64}_16x16x4): X[p][nu] = sum_mu phi0[p][mu] D[mu][nu];
//     the epilogue forms rho_c[p] = sum_nu X[p][nu] phi_c[p][nu] from the
//     accumulators, so X never goes to LDS (wave w owns point block w % PB and
//     every (4 / PB)-th nu block; the 4 / PB partial sums meet in LDS);
//   pointwise B3LYP; Y[p][nu] = c0 phi0 + sum_c c_c phi_c;
//   GEMM 2: V[mu][nu] += sum_p phi0[p][mu] Y[p][nu], 16 x 16 blocks of V held
//     in accumulator registers across all tiles (block wave + 4u).
// Operand layout of v_mfma_*_16x16x4 (lane l): A[l % 16][l / 16],
// B[l / 16][l % 16]; accumulator register j holds D[r][l % 16] with
// r = 4 (l / 16) + j for f32 and r = l / 16 + 4 j for f64.  XS = 4 mod 16 keeps the GEMM-1
// A-operand and epilogue reads of Phi free of bank conflicts; D and Y use row
// stride XD = 16 mod 64 for the same reason.
template <typename T>
struct XMTraits;
template <>
struct XMTraits<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static constexpr int TILE = 64;
  static __device__ inline v4 mfma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  // row of accumulator register j in lane group lg = lane / 16: row(lg, 0) + j * RSTEP
  static __device__ inline int row(int lg, int j) { return 4 * lg + j; }
  static constexpr int RSTEP = 1;
};
template <>
struct XMTraits<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static constexpr int TILE = 32;
  static __device__ inline v4 mfma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lg, int j) { return lg + 4 * j; }
  static constexpr int RSTEP = 4;
};
constexpr int XM_MAXN = 80;  // nao limit of the MFMA path

__host__ __device__ inline int xm_np(int N) { return (N + 15) & ~15; }

// Shapes of the tile pipeline for XT points and nao padded to NP (a multiple
// of 16 up to 80), all compile-time so that every LDS offset in the GEMM and
// epilogue loops is an immediate.
template <int XT, int NP>
struct XTile {
  static constexpr int PB = XT / 16, NSLOT = 4 / PB;      // GEMM-1: point blocks, column-block slots
  static constexpr int NB = NP / 16, NJ = (NB + NSLOT - 1) / NSLOT;
  static constexpr int XS = NP + 4;                        // Phi row stride, 4 mod 16
  static constexpr int XD = NP <= 16 ? 16 : 80;            // D and Y row stride, 16 mod 64
  static constexpr int NVB = NB * NB, NBW = (NVB + 3) / 4; // V blocks, per wave
  static constexpr size_t lds_elems = (size_t)4 * XT * XS + (size_t)XT * XD + (size_t)NP * XD + 256 + 4 * XT;
};

// FP32 with 32-point tiles and nao <= 64 keeps GEMM 1's D operand in registers
// (NJ * NP / 4 values per lane: D does not change between tiles) instead of
// LDS, which lowers the LDS of a workgroup from 67 KB to 46 KB at NP = 64, so
// three workgroups share a CU instead of two.  FP64 and wider shapes would need
// more registers than they have spare and keep D in LDS.
template <typename T, int XT, int NP>
struct XDReg {
  static constexpr bool on = sizeof(T) == 4 && XT == 32 && NP <= 64;
  static constexpr size_t lds_elems = XTile<XT, NP>::lds_elems - (on ? (size_t)NP * XTile<XT, NP>::XD : 0);
};

template <typename T, int XT, int NP>
using XVacc = typename XMTraits<T>::v4[XTile<XT, NP>::NBW];

// Staging of a tile's AO values and gradients into Phi[c][p][mu] (row stride
// XS), mu < N; padding columns are left as they are.
// Stored form: read from ao[c * sc + mu * smu + p * sp] in HBM.  Each wave
// moves one (c, mu) row of XT points per 64 / XT rows and instruction (lanes
// walk points, coalesced), eight rows' loads in flight before their LDS
// stores; c = row / N comes from three compares, not a division.
template <typename T, int XT, int NP>
struct XcStageStored {
  const T* ao;
  int sc, smu, sp, g, N;
  __device__ inline void load(int) {}
  __device__ inline void operator()(T* Phi, int p0, int /*p0_next*/) {
    constexpr int XS = XTile<XT, NP>::XS;
    constexpr int RPW = 64 / XT;  // rows per wave instruction
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int p = lane % XT, sub = lane / XT, nrow = 4 * N, rstride = 4 * RPW;
    const int pt = p0 + p;
    const bool in = pt < g;
    const int ptoff = pt * sp;
    for (int base = wave * RPW + sub; base < nrow; base += 8 * rstride) {
      T v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int row = base + u * rstride;
        const int c = (row >= N) + (row >= 2 * N) + (row >= 3 * N), mu = row - c * N;
        v[u] = (row < nrow && in) ? ao[c * sc + mu * smu + ptoff] : T(0);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int row = base + u * rstride;
        const int c = (row >= N) + (row >= 2 * N) + (row >= 3 * N), mu = row - c * N;
        if (row < nrow) Phi[(c * XT + p) * XS + mu] = v[u];
      }
    }
  }
};

// One molecule's shells in LDS for the direct form, as records: an s shell,
// a p shell, or an s shell and the p shell after it on the same atom with the
// same exponents (STO-3G's and 6-31G's sp shells), which then share their
// exponentials.  Records are ordered s, sp, p, so the two records a wave
// evaluates at once mostly take the same branch.
//   hdr[4 r]:  centre x, y, z and the squared distance beyond which every
//              primitive's exponential is below 2^-48 (such values are written as 0);
//   info[r]:   mu | kind << 8 | nprim << 12, kind 0 = s, 1 = p, 2 = sp (s at
//              mu, p at mu + 1 .. mu + 3), mu = the record's first AO;
//   prim[r kp + k]: {e log2(e), e, c_s, c_p}, the contraction coefficients
//              times the angular normalisation (c_s = 0 for p, c_p = 0 for s).
template <typename T>
struct alignas(16) XcPrim {
  T a, e, cs, cp;
};

template <typename T>
struct XcShells {
  double* hdr;
  int* info;
  XcPrim<T>* prim;
  int nrec, kp;
};

__host__ __device__ inline size_t xc_shell_prim_off(int nsh_max) {
  return ((size_t)nsh_max * (4 * sizeof(double) + 3 * sizeof(int)) + 31) & ~(size_t)31;
}

__host__ __device__ inline size_t xc_shell_bytes(int nsh_max, int kp, size_t es) {
  return xc_shell_prim_off(nsh_max) + (size_t)nsh_max * kp * 4 * es;
}

// Builds the table at `base` (32-byte aligned) from the molecule's shells
// bas[bas0 .. bas0 + nsh) (the s and p shells eval_gto evaluates,
// nanoDFT.py:219-220).  Ends with a barrier.  A shell with more than kp
// primitives gets NaN coefficients, so its molecule's result is NaN.
template <typename T>
__device__ inline XcShells<T> xc_build_shells(unsigned char* base, int nsh_max, int kp, int bas0, int nsh,
                                              const int* __restrict__ atm, const int* __restrict__ bas,
                                              const double* __restrict__ env) {
  XcShells<T> t;
  t.hdr = reinterpret_cast<double*>(base);
  t.info = reinterpret_cast<int*>(t.hdr + 4 * nsh_max);
  int* lsh = t.info + nsh_max;     // l of each shell
  int* okey = lsh + nsh_max;       // record order key: 0 s, 1 sp, 2 p, 3 none (p merged into the s before it)
  t.prim = reinterpret_cast<XcPrim<T>*>(base + xc_shell_prim_off(nsh_max));
  t.kp = kp;
  const int tid = threadIdx.x;
  for (int i = tid; i < nsh; i += blockDim.x) lsh[i] = bas[(bas0 + i) * BAS_SLOTS + 1];
  __syncthreads();
  // s shell i and p shell i + 1 merge when they share atom and exponents
  auto merges = [&](int i) {
    if (i + 1 >= nsh || lsh[i] != 0 || lsh[i + 1] != 1) return false;
    const int* b = bas + (bas0 + i) * BAS_SLOTS;
    const int* q = b + BAS_SLOTS;
    if (b[0] != q[0] || b[2] != q[2]) return false;
    for (int k = 0; k < b[2]; ++k)
      if (env[b[5] + k] != env[q[5] + k]) return false;
    return true;
  };
  for (int i = tid; i < nsh; i += blockDim.x) {
    const bool into_prev = i > 0 && merges(i - 1);
    okey[i] = into_prev ? 3 : (merges(i) ? 1 : (lsh[i] == 0 ? 0 : 2));
  }
  __syncthreads();
  int nrec = 0;
  for (int j = 0; j < nsh; ++j) nrec += okey[j] < 3;
  t.nrec = nrec;
  for (int i = tid; i < nsh; i += blockDim.x) {
    const int key = okey[i];
    if (key == 3) continue;
    int mu = 0, pos = 0;
    for (int j = 0; j < nsh; ++j) {
      if (j < i) mu += 2 * lsh[j] + 1;
      pos += okey[j] < key || (okey[j] == key && j < i);
    }
    const int* b = bas + (bas0 + i) * BAS_SLOTS;
    const int* q = b + BAS_SLOTS;  // the p shell of an sp record
    const double* R = env + atm[b[0] * ATM_SLOTS + 1];
    const int np = b[2];
    const int npk = np < kp ? np : kp;
    const int kind = key == 0 ? 0 : (key == 1 ? 2 : 1);
    double amin = 1e300;
    for (int k = 0; k < npk; ++k) {
      const double e = env[b[5] + k];
      const double cs = kind == 1 ? 0.0 : ANG_S * env[b[6] + k];
      const double cp = kind == 0 ? 0.0 : ANG_P * env[(kind == 2 ? q : b)[6] + k];
      XcPrim<T> pr;
      pr.a = (T)(e * 1.4426950408889634);  // log2(e)
      pr.e = (T)e;
      pr.cs = np > kp ? T(NAN) : (T)cs;
      pr.cp = np > kp ? T(NAN) : (T)cp;
      t.prim[pos * kp + k] = pr;
      amin = fmin(amin, e * 1.4426950408889634);
    }
    // unused slots contribute 2^0 times zero coefficients
    for (int k = npk; k < kp; ++k) t.prim[pos * kp + k] = XcPrim<T>{T(0), T(0), T(0), T(0)};
    t.hdr[4 * pos] = R[0];
    t.hdr[4 * pos + 1] = R[1];
    t.hdr[4 * pos + 2] = R[2];
    t.hdr[4 * pos + 3] = 48.0 / amin;
    t.info[pos] = mu | (kind << 8) | (npk << 12);
  }
  __syncthreads();
  return t;
}

// Direct form: the AO values and gradients of the tile's points evaluated
// from the shell records (the values pscf_batch_ao stores).  Thread t takes
// point t % XT of the tile (XC_THREADS is a multiple of XT, so a thread keeps
// its point) and records t / XT, t / XT + XC_THREADS / XT, ...; the point's
// coordinates for the next tile are loaded at the end of each stage, so their
// latency hides behind the tile's GEMMs.  Differences to the centres are
// formed in FP64 and rounded to T; the exponentials are 2^(-a r^2) with
// a = e log2(e) (one v_exp_f32 in FP32).
template <typename T, int XT, int NP>
struct XcStageDirect {
  XcShells<T> sh;
  const double* coords;  // the molecule's points, [g][3]
  int g;
  double px, py, pz;
  __device__ inline void load(int p0) {
    const int pt = p0 + (int)(threadIdx.x % XT);
    const bool in = pt < g;
    px = in ? coords[3 * pt] : 0.0;
    py = in ? coords[3 * pt + 1] : 0.0;
    pz = in ? coords[3 * pt + 2] : 0.0;
  }
  __device__ static inline T exp2n(T x) {
    if constexpr (sizeof(T) == 4)
      return __builtin_amdgcn_exp2f(x);
    else
      return exp2(x);
  }
  __device__ inline void operator()(T* Phi, int p0, int p0_next) {
    constexpr int XS = XTile<XT, NP>::XS;
    const int tid = threadIdx.x, p = tid % XT;
    const bool in = p0 + p < g;
    const double x = px, y = py, z = pz;
    load(p0_next);
    for (int r = tid / XT; r < sh.nrec; r += XC_THREADS / XT) {
      const int info = sh.info[r];
      const int mu = info & 255, kind = (info >> 8) & 15, np = info >> 12;
      const double* h = sh.hdr + 4 * r;
      const T dx = (T)(x - h[0]), dy = (T)(y - h[1]), dz = (T)(z - h[2]);
      const T r2 = dx * dx + dy * dy + dz * dz;
      // the first three primitive slots are evaluated unconditionally (kp >=
      // 3, unused slots hold zero coefficients), so their loads issue with the
      // header's and no lane branches on the primitive count
      XcPrim<T> q3[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) q3[k] = sh.prim[r * sh.kp + k];
      T rs = T(0), ds = T(0), rp = T(0), dp = T(0);
      auto add = [&](const XcPrim<T>& q) {
        const T ex = exp2n(-q.a * r2);
        const T te = q.e * ex;
        rs += q.cs * ex;
        ds += q.cs * te;
        rp += q.cp * ex;
        dp += q.cp * te;
      };
      if (in && (double)r2 <= h[3]) {
#pragma unroll
        for (int k = 0; k < 3; ++k) add(q3[k]);
        for (int k = 3; k < np; ++k) add(sh.prim[r * sh.kp + k]);
      }
      ds *= T(-2);
      dp *= T(-2);
      T* qs = Phi + p * XS + mu;
      if (kind != 1) {
        qs[0] = rs;
        qs[XT * XS] = ds * dx;
        qs[2 * XT * XS] = ds * dy;
        qs[3 * XT * XS] = ds * dz;
      }
      if (kind != 0) {
        T* qp = qs + (kind == 2 ? 1 : 0);
        const T d[3] = {dx, dy, dz};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          qp[a] = d[a] * rp;
#pragma unroll
          for (int c = 0; c < 3; ++c) qp[(1 + c) * XT * XS + a] = (a == c ? rp : T(0)) + d[a] * d[c] * dp;
        }
      }
    }
  }
};

template <typename T, int XT, int NP, class Stage>
__device__ inline void xc_tiles(const XCArgs<T>& a, unsigned char* smem, int tile0, int tstride, int tile_end,
                                Stage& stage, XVacc<T, XT, NP>& vacc, double& e_loc) {
  typedef typename XMTraits<T>::v4 v4;
  typedef XTile<XT, NP> L;
  constexpr int PB = L::PB, NSLOT = L::NSLOT, NB = L::NB, XS = L::XS, XD = L::XD, NVB = L::NVB, NBW = L::NBW;
  constexpr int RS = XMTraits<T>::RSTEP;
  const int N = a.N;
  constexpr bool DREG = XDReg<T, XT, NP>::on;
  constexpr int NK = NP / 4;
  T* Phi = reinterpret_cast<T*>(smem);        // [4][XT][XS]
  T* Y = Phi + 4 * XT * XS;                   // [XT][XD]
  T* Dm = Y + XT * XD;                        // [NP][XD] (not when DREG)
  T* rpart = Dm + (DREG ? 0 : NP * XD);       // [NSLOT][4][XT]
  T* coef = rpart + 256;                      // [4][XT]
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lg = lane >> 4;
  const int pb = wave % PB, slot = wave / PB;
  // GEMM 1's B operand of this lane: D[k0 + lg][column block of jb, li]
  T dreg[DREG ? L::NJ * NK : 1];
  __syncthreads();
  if (DREG) {
#pragma unroll
    for (int jb = 0; jb < L::NJ; ++jb) {
      const int cbk = slot + jb * NSLOT < NB ? slot + jb * NSLOT : NB - 1;  // as in GEMM 1
      const int col = cbk * 16 + li;
#pragma unroll
      for (int kk = 0; kk < NK; ++kk) {
        const int row = kk * 4 + lg;
        dreg[DREG ? jb * NK + kk : 0] = (row < N && col < N) ? a.dm[row * N + col] : T(0);
      }
    }
  } else {
    for (int x = tid; x < NP * NP; x += blockDim.x) {
      const int i = x / NP, j = x - i * NP;
      Dm[i * XD + j] = (i < N && j < N) ? a.dm[i * N + j] : T(0);
    }
  }
  // padding columns nu in [N, NP) of Phi stay zero for every tile
  for (int x = tid; x < 4 * XT * (NP - N); x += blockDim.x) {
    const int row = x / (NP - N);
    Phi[row * XS + N + x - row * (NP - N)] = T(0);
  }
#pragma unroll
  for (int u = 0; u < NBW; ++u) vacc[u] = (v4){T(0), T(0), T(0), T(0)};
  // per-lane LDS offsets, fixed for all tiles
  const int a1 = (pb * 16 + li) * XS + lg;                           // GEMM 1 A: + k0
  const int b1 = lg * XD + slot * 16 + li;                           // GEMM 1 B: + k0 XD + jb NSLOT 16
  const int e1 = (pb * 16 + XMTraits<T>::row(lg, 0)) * XS + slot * 16 + li;  // epilogue
  const int a2 = lg * XS + li, b2 = lg * XD + li;                    // GEMM 2: + block and k0 offsets
  // Every load above (D, the first tile's coordinates) completes here.  The
  // loop then carries only its own prefetches (the next tile's coordinates,
  // this tile's weights), each waited for where it is used, instead of the
  // compiler merging the loop entry's pending loads into a full wait inside
  // every tile.
  __builtin_amdgcn_s_waitcnt(0);
  for (int tile = tile0; tile < tile_end; tile += tstride) {
    const int p0 = tile * XT;
    // this tile's weight for the pointwise step, loaded before the staging
    const T wpre = (tid < XT && p0 + tid < a.g) ? a.w[p0 + tid] : T(0);
    __syncthreads();
    stage(Phi, p0, p0 + tstride * XT);
    __syncthreads();
    // GEMM 1 and the density epilogue: wave (pb, slot) takes point block pb
    // and the column blocks slot, slot + NSLOT, ...; its 16 partial sums
    // rho_c of 4 points are reduced over the 16 lanes of each DPP row at once,
    // after which every lane holds one of them.
    {
      T part[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) part[k] = T(0);
#pragma unroll
      for (int jb = 0; jb < L::NJ; ++jb) {
        // column block slot + jb NSLOT; past NB the wave repeats block NB - 1
        // with its sums scaled by zero (no branch, so the accumulators stay
        // in place)
        const bool on = slot + jb * NSLOT < NB;
        const int cb = on ? jb * NSLOT * 16 : (NB - 1 - slot) * 16;
        const T scale = on ? T(1) : T(0);
        v4 acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int k0 = 0; k0 < NP; k0 += 4)
          acc = XMTraits<T>::mfma(Phi[a1 + k0], DREG ? dreg[DREG ? jb * NK + k0 / 4 : 0] : Dm[b1 + k0 * XD + cb], acc);
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] *= scale;
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
          for (int j = 0; j < 4; ++j) part[c * 4 + j] += acc[j] * Phi[e1 + c * XT * XS + j * RS * XS + cb];
      }
      const T tot = transpose_reduce16_row(part, lane);
      const int c = li >> 2, j = li & 3;
      rpart[(slot * 4 + c) * XT + pb * 16 + XMTraits<T>::row(lg, j)] = tot;
    }
    __syncthreads();
    // pointwise B3LYP: wave w < 3 evaluates part w of the functional
    // (Slater + VWN, B88, LYP) at the tile's points (lane = point); the parts
    // meet in LDS (Y's space, free until Y is formed) and the first XT
    // threads, which hold rho and its gradient, combine them
    T rc[4];
    if (wave < 3 && lane < XT) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        T v = T(0);
#pragma unroll
        for (int q = 0; q < NSLOT; ++q) v += rpart[(q * 4 + c) * XT + lane];
        rc[c] = v;
      }
      T rho, gam, e = T(0), vr = T(0), vn = T(0);
      b3lyp_inputs<T>(rc[0], rc[1], rc[2], rc[3], a.eps, rho, gam);
      if (wave == 0)
        b3lyp_part<T, 0>(rho, gam, e, vr, vn);
      else if (wave == 1)
        b3lyp_part<T, 1>(rho, gam, e, vr, vn);
      else
        b3lyp_part<T, 2>(rho, gam, e, vr, vn);
      Y[(3 * wave) * XT + lane] = e;
      Y[(3 * wave + 1) * XT + lane] = vr;
      Y[(3 * wave + 2) * XT + lane] = vn;
    }
    __syncthreads();
    if (tid < XT) {
      const int pt = p0 + tid;
      T c0 = T(0), c1 = T(0), c2 = T(0), c3 = T(0);
      if (pt < a.g) {
        const T e = Y[tid] + Y[3 * XT + tid] + Y[6 * XT + tid];
        const T vr = Y[XT + tid] + Y[4 * XT + tid] + Y[7 * XT + tid];
        const T vn = Y[2 * XT + tid] + Y[5 * XT + tid] + Y[8 * XT + tid];
        const T wt = wpre;
        e_loc += (double)rc[0] * (double)wt * (double)e;
        c0 = T(0.5) * vr * wt;
        const T f = T(4) * vn * wt;
        c1 = f * rc[1];
        c2 = f * rc[2];
        c3 = f * rc[3];
      }
      coef[tid] = c0;
      coef[XT + tid] = c1;
      coef[2 * XT + tid] = c2;
      coef[3 * XT + tid] = c3;
    }
    __syncthreads();
    // Y[p][nu] = c0 phi0 + sum_c c_c phi_c (XT NP is a multiple of the 256 threads)
#pragma unroll
    for (int it = 0; it < XT * NP / XC_THREADS; ++it) {
      const int x = tid + it * XC_THREADS;
      const int p = x / NP, nu = x - p * NP;
      Y[p * XD + nu] = coef[p] * Phi[p * XS + nu] + coef[XT + p] * Phi[(XT + p) * XS + nu] +
                       coef[2 * XT + p] * Phi[(2 * XT + p) * XS + nu] + coef[3 * XT + p] * Phi[(3 * XT + p) * XS + nu];
    }
    __syncthreads();
    // GEMM 2: V[mu][nu] += sum_p phi0[p][mu] Y[p][nu], blocks wave + 4u (a
    // block index past NVB repeats the last block, whose copy is not stored)
#pragma unroll
    for (int k0 = 0; k0 < XT; k0 += 4) {
#pragma unroll
      for (int u = 0; u < NBW; ++u) {
        const int vb = min(wave + 4 * u, NVB - 1);
        const int mb = vb / NB, nb = vb - mb * NB;
        vacc[u] = XMTraits<T>::mfma(Phi[a2 + k0 * XS + mb * 16], Y[b2 + k0 * XD + nb * 16], vacc[u]);
      }
    }
  }
}

// V blocks of this wave (rows/cols < N) into dst with row stride ld.
template <typename T, int XT, int NP>
__device__ inline void xm_store_v(const XVacc<T, XT, NP>& vacc, int N, T* dst, int ld) {
  constexpr int NB = NP / 16, NBW = XTile<XT, NP>::NBW;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, lg = lane >> 4;
#pragma unroll
  for (int u = 0; u < NBW; ++u) {
    const int vb = wave + 4 * u;
    if (vb < NB * NB) {
      const int mb = vb / NB, nb = vb - mb * NB;
      const int col = nb * 16 + li;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = mb * 16 + XMTraits<T>::row(lg, j);
        if (row < N && col < N) dst[row * ld + col] = vacc[u][j];
      }
    }
  }
}

__device__ inline double xc_block_energy(double e_loc, double* red) {
  for (int off = 32; off > 0; off >>= 1) e_loc += __shfl_xor(e_loc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e_loc;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < XC_THREADS / 64; ++k) s += red[k];
  return s;
}

template <typename T, int NP>
__global__ void __launch_bounds__(XC_THREADS) xc_fused_kernel(XCArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int XT = XMTraits<T>::TILE;
  double e_loc = 0.0;
  XVacc<T, XT, NP> vacc;
  // AO strides as 32-bit element offsets (the host checks that the AO block
  // is below 2^31 elements)
  XcStageStored<T, XT, NP> stage{a.ao, (int)a.sc, (int)a.smu, (int)a.sp, a.g, a.N};
  xc_tiles<T, XT, NP>(a, smem, blockIdx.x, gridDim.x, (a.g + XT - 1) / XT, stage, vacc, e_loc);
  double* red = reinterpret_cast<double*>(smem);
  if constexpr (sizeof(T) == 8) {
    if (a.vout != nullptr) {
      // this workgroup's V + V^T and energy into the outputs (exact atomics)
      __syncthreads();
      const double e = xc_block_energy(e_loc, red);
      if (threadIdx.x == 0) atomicAdd(a.eout, quant40(e));
      __syncthreads();
      T* Vl = reinterpret_cast<T*>(smem);
      xm_store_v<T, XT, NP>(vacc, a.N, Vl, a.N);
      __syncthreads();
      const int NN = a.N * a.N;
      for (int x = threadIdx.x; x < NN; x += blockDim.x) {
        const int i = x / a.N, j = x - i * a.N;
        atomicAdd(a.vout + x, quant40(Vl[x] + Vl[j * a.N + i]));
      }
      return;
    }
  }
  xm_store_v<T, XT, NP>(vacc, a.N, a.part_v + (int64_t)blockIdx.x * a.N * a.N, a.N);
  __syncthreads();
  const double e = xc_block_energy(e_loc, red);
  if (threadIdx.x == 0) a.part_e[blockIdx.x] = e;
}

// Batched: workgroup (blockIdx.x, blockIdx.y) takes molecule order[blockIdx.x]
// (nao <= max_nao <= NP) and skips converged molecules.  A molecule's grid is
// cut into chunks of XC_CHUNK_PTS points; part blockIdx.y runs chunks
// blockIdx.y, blockIdx.y + gridDim.y, ..., each from zeroed accumulators, and
// stores the chunk's V and E_xc partials in slot (chunk, blockIdx.x) of part_v
// ([slot][max_nao^2], row stride nao) and part_e; xc_batch_reduce_kernel adds
// the chunks in order.  Each chunk's sums cover the same points in the same
// order whatever the number of parts, so the result does not depend on the
// split.  A molecule with a single chunk writes V + V^T and E_xc directly.
// Slots beyond the part buffers' capacities are not written; the reduction
// marks such a molecule with NaN.
constexpr int XC_CHUNK_PTS = PSCF_XC_CHUNK_PTS;

__host__ __device__ inline bool xc_parts_fit(int nchunk, int nmol, int max_nao, int64_t part_v_elems,
                                             int64_t part_e_elems) {
  return (int64_t)nchunk * nmol * max_nao * max_nao <= part_v_elems && (int64_t)nchunk * nmol <= part_e_elems;
}

// Direct form of the batched kernel: AO values from the points' coordinates
// and the molecules' shells (bas/env, as pscf_batch_ao reads them) instead of
// a stored AO block; the shell table of nsh_max shells with kp primitive slots
// sits at byte tab_off of the LDS.
struct XcDirect {
  const double* coords;
  const int* atm;
  const int* bas;
  const double* env;
  int nsh_max, kp, tab_off;
};

// A molecule the launch's shapes do not cover (nao above max_nao, a 32-bit
// AO offset overflow, or in the direct form more shells than the table holds).
__device__ inline bool xc_mol_outside(const BatchMol& Mo, int max_nao, int nsh_max) {
  return Mo.nao > max_nao || (int64_t)4 * Mo.nao * Mo.ngrid >= ((int64_t)1 << 31) ||
         (nsh_max > 0 && Mo.nbas > nsh_max);
}

// FP32 with 32-point tiles: three workgroups per CU up to nao 64 (their LDS
// fits), two above; the register cap this sets costs a few spilled values
// outside the GEMM loops and runs faster than fewer resident workgroups.
template <typename T, int XT, int NP>
constexpr int xc_batch_min_blocks() {
  return (sizeof(T) == 4 && XT == 32) ? (NP <= 64 ? 3 : 2) : 1;
}

template <typename T, int XT, int NP, bool DIRECT>
__global__ void __launch_bounds__(XC_THREADS, (xc_batch_min_blocks<T, XT, NP>())) xc_batch_kernel(const BatchMol* __restrict__ mols,
                                                              const int* __restrict__ order, int max_nao,
                                                              const double* __restrict__ ctl, const T* __restrict__ ao,
                                                              XcDirect da, const T* __restrict__ dm,
                                                              const T* __restrict__ w, T eps, T* __restrict__ vxc,
                                                              double* __restrict__ exc, T* __restrict__ part_v,
                                                              double* __restrict__ part_e, int64_t part_v_elems,
                                                              int64_t part_e_elems) {
  static_assert(XC_CHUNK_PTS % XT == 0, "a chunk is a whole number of tiles");
  extern __shared__ __align__(16) unsigned char smem[];
  const int m = order[blockIdx.x];
  if (ctl != nullptr && ctl[m * CTL + 1] == 0.0) return;
  const BatchMol Mo = mols[m];
  const int N = (int)Mo.nao;
  const int part = blockIdx.y, nparts = gridDim.y;
  // LDS and tiles were sized for max_nao (and nsh_max shells), offsets are
  // 32-bit: flag a molecule outside those bounds
  if (xc_mol_outside(Mo, max_nao, DIRECT ? da.nsh_max : 0)) {
    if (part == 0) {
      for (int x = threadIdx.x; x < N * N; x += blockDim.x) vxc[Mo.mat0 + x] = T(NAN);
      if (threadIdx.x == 0) exc[m] = NAN;
    }
    return;
  }
  XCArgs<T> a;
  a.vout = nullptr;
  a.eout = nullptr;
  a.N = N;
  a.g = (int)Mo.ngrid;
  a.ao = DIRECT ? nullptr : ao + Mo.ao0;
  a.sc = (int64_t)a.N * a.g;
  a.smu = a.g;
  a.sp = 1;
  a.dm = dm + Mo.mat0;
  a.w = w + Mo.grid0;
  a.eps = eps;
  constexpr int TPC = XC_CHUNK_PTS / XT;  // tiles per chunk
  const int ntiles = (a.g + XT - 1) / XT;
  const int nchunk = (a.g + XC_CHUNK_PTS - 1) / XC_CHUNK_PTS;
  double* red = reinterpret_cast<double*>(smem + sizeof(T) * N * N);
  XVacc<T, XT, NP> vacc;
  auto run = [&](auto& stage) {
    if (nchunk <= 1) {
      if (part != 0) return;
      double e_loc = 0.0;
      stage.load(0);
      xc_tiles<T, XT, NP>(a, smem, 0, 1, ntiles, stage, vacc, e_loc);
      __syncthreads();
      T* Vl = reinterpret_cast<T*>(smem);
      xm_store_v<T, XT, NP>(vacc, N, Vl, N);
      __syncthreads();
      T* out = vxc + Mo.mat0;
      for (int x = threadIdx.x; x < N * N; x += blockDim.x) {
        const int i = x / N, j = x - i * N;
        out[x] = Vl[x] + Vl[j * N + i];
      }
      __syncthreads();
      const double e = xc_block_energy(e_loc, red);
      if (threadIdx.x == 0) exc[m] = e;
      return;
    }
    if (!xc_parts_fit(nchunk, gridDim.x, max_nao, part_v_elems, part_e_elems)) return;  // marked by the reduction
    for (int c = part; c < nchunk; c += nparts) {
      double e_loc = 0.0;
      stage.load(c * XC_CHUNK_PTS);
      xc_tiles<T, XT, NP>(a, smem, c * TPC, 1, min(ntiles, (c + 1) * TPC), stage, vacc, e_loc);
      __syncthreads();
      const int64_t slot = (int64_t)c * gridDim.x + blockIdx.x;
      xm_store_v<T, XT, NP>(vacc, N, part_v + slot * max_nao * max_nao, N);
      const double e = xc_block_energy(e_loc, red);
      if (threadIdx.x == 0) part_e[slot] = e;
    }
  };
  if constexpr (DIRECT) {
    XcStageDirect<T, XT, NP> stage;
    stage.sh = xc_build_shells<T>(smem + da.tab_off, da.nsh_max, da.kp, (int)Mo.bas0, (int)Mo.nbas, da.atm, da.bas,
                                  da.env);
    stage.coords = da.coords + 3 * Mo.grid0;
    stage.g = a.g;
    run(stage);
  } else {
    XcStageStored<T, XT, NP> stage{a.ao, (int)a.sc, (int)a.smu, (int)a.sp, a.g, a.N};
    run(stage);
  }
}

// Sums the chunks of xc_batch_kernel in order: V_xc = sum_c (V_c + V_c^T),
// E_xc = sum_c e_c, one workgroup per active molecule with more than one chunk.
template <typename T>
__global__ void __launch_bounds__(256) xc_batch_reduce_kernel(const BatchMol* __restrict__ mols,
                                                              const int* __restrict__ order, int max_nao,
                                                              const double* __restrict__ ctl,
                                                              const T* __restrict__ part_v,
                                                              const double* __restrict__ part_e, int64_t part_v_elems,
                                                              int64_t part_e_elems, T* __restrict__ vxc,
                                                              double* __restrict__ exc, int nsh_max) {
  const int m = order[blockIdx.x];
  if (ctl != nullptr && ctl[m * CTL + 1] == 0.0) return;
  const BatchMol Mo = mols[m];
  const int N = (int)Mo.nao;
  if (xc_mol_outside(Mo, max_nao, nsh_max)) return;  // flagged by xc_batch_kernel
  const int nchunk = (int)((Mo.ngrid + XC_CHUNK_PTS - 1) / XC_CHUNK_PTS);
  if (nchunk <= 1) return;  // written by xc_batch_kernel
  if (!xc_parts_fit(nchunk, gridDim.x, max_nao, part_v_elems, part_e_elems)) {
    for (int x = threadIdx.x; x < N * N; x += blockDim.x) vxc[Mo.mat0 + x] = T(NAN);
    if (threadIdx.x == 0) exc[m] = NAN;
    return;
  }
  // chunk sums element by element (coalesced), in chunk order, into LDS;
  // then V + V^T from LDS
  extern __shared__ __align__(16) unsigned char smem_r[];
  T* Vs = reinterpret_cast<T*>(smem_r);
  for (int x = threadIdx.x; x < N * N; x += blockDim.x) {
    T v = T(0);
    for (int c = 0; c < nchunk; ++c) v += part_v[((int64_t)c * gridDim.x + blockIdx.x) * max_nao * max_nao + x];
    Vs[x] = v;
  }
  __syncthreads();
  for (int x = threadIdx.x; x < N * N; x += blockDim.x) {
    const int i = x / N, j = x - i * N;
    vxc[Mo.mat0 + x] = Vs[x] + Vs[j * N + i];
  }
  if (threadIdx.x == 0) {
    double e = 0.0;
    for (int c = 0; c < nchunk; ++c) e += part_e[(int64_t)c * gridDim.x + blockIdx.x];
    exc[m] = e;
  }
}

// Grid points of one atom's template: chunk (atom, first point) per workgroup.
__global__ void __launch_bounds__(256) grid_expand_kernel(const int* __restrict__ work, const int* __restrict__ atom_elem,
                                                          const int64_t* __restrict__ atom_pt0,
                                                          const int* __restrict__ atom_mol,
                                                          const int* __restrict__ tmpl_off,
                                                          const int* __restrict__ tmpl_n,
                                                          const double* __restrict__ tmpl, const int* __restrict__ atm,
                                                          const double* __restrict__ env, double* __restrict__ coords,
                                                          double* __restrict__ vol, int* __restrict__ owner,
                                                          int* __restrict__ ptmol) {
  const int a = work[2 * blockIdx.x];
  const int i = work[2 * blockIdx.x + 1] + threadIdx.x;
  const int e = atom_elem[a];
  if (i >= tmpl_n[e]) return;
  const double* t = tmpl + 4 * ((int64_t)tmpl_off[e] + i);
  const double* R = env + atm[a * ATM_SLOTS + 1];
  const int64_t g = atom_pt0[a] + i;
  coords[3 * g] = R[0] + t[0];
  coords[3 * g + 1] = R[1] + t[1];
  coords[3 * g + 2] = R[2] + t[2];
  vol[g] = t[3];
  owner[g] = a;
  ptmol[g] = atom_mol[a];
}

// Bragg-Slater radii (Bohr), H..Ne, for the Treutler adjustment.
__constant__ double c_bragg[11] = {0.0,
                                   0.35 / 0.52917721092, 1.40 / 0.52917721092, 1.45 / 0.52917721092,
                                   1.05 / 0.52917721092, 0.85 / 0.52917721092, 0.70 / 0.52917721092,
                                   0.65 / 0.52917721092, 0.60 / 0.52917721092, 0.50 / 0.52917721092,
                                   1.50 / 0.52917721092};

// Becke weights of all points of all molecules; each point's cell function runs
// over its own molecule's atoms only.  T: arithmetic (FP64 or FP32).
// Becke's smoothing step applied three times to nu (an odd function of nu).
template <typename T>
__device__ inline T becke_f3(T g) {
  g = (T(3) - g * g) * g * T(0.5);
  g = (T(3) - g * g) * g * T(0.5);
  return (T(3) - g * g) * g * T(0.5);
}

// Molecules of at most BECKE_FAST_MAXA atoms: each point's atom distances in
// registers, computed once (instead of once per ordered pair), and every
// unordered pair visited once for both its ordered cell functions, each with
// its own adjustment a_ij / a_ji from the table.
constexpr int BECKE_FAST_MAXA = 32;

__global__ void __launch_bounds__(256) becke_fast_kernel(const double* __restrict__ coords,
                                                         const int* __restrict__ owner,
                                                         const double* __restrict__ vol, int g,
                                                         const double* __restrict__ atom_xyz,
                                                         const double* __restrict__ adj, int natm,
                                                         double* __restrict__ w) {
  constexpr int MA = BECKE_FAST_MAXA;
  __shared__ double s_xyz[3 * MA], s_inv[MA * MA], s_a[MA * MA];
  for (int i = threadIdx.x; i < 3 * natm; i += blockDim.x) s_xyz[i] = atom_xyz[i];
  for (int q = threadIdx.x; q < natm * natm; q += blockDim.x) {
    const int p = q / natm, r = q - p * natm;
    const double dx = atom_xyz[3 * p] - atom_xyz[3 * r], dy = atom_xyz[3 * p + 1] - atom_xyz[3 * r + 1],
                 dz = atom_xyz[3 * p + 2] - atom_xyz[3 * r + 2];
    s_inv[p * MA + r] = p == r ? 0.0 : 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    s_a[p * MA + r] = adj[q];
  }
  __syncthreads();
  const int pt = blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= g) return;
  const double x = coords[3 * pt], y = coords[3 * pt + 1], z = coords[3 * pt + 2];
  double r[MA], P[MA];
#pragma unroll
  for (int i = 0; i < MA; ++i) {
    P[i] = 1.0;
    r[i] = 0.0;
    if (i < natm) {
      const double dx = x - s_xyz[3 * i], dy = y - s_xyz[3 * i + 1], dz = z - s_xyz[3 * i + 2];
      r[i] = sqrt(dx * dx + dy * dy + dz * dz);
    }
  }
  // (a uniform branch per i rather than a cut-short loop, so both loops stay
  // unrolled and r, P in registers)
#pragma unroll
  for (int i = 1; i < MA; ++i) {
    if (i < natm) {
#pragma unroll
      for (int j = 0; j < i; ++j) {
        const double mu = (r[i] - r[j]) * s_inv[i * MA + j];
        const double om = 1.0 - mu * mu;
        P[i] *= 0.5 * (1.0 - becke_f3(mu + s_a[i * MA + j] * om));
        P[j] *= 0.5 * (1.0 - becke_f3(-mu + s_a[j * MA + i] * om));
      }
    }
  }
  const int own = owner[pt];
  double psum = 0.0, pown = 0.0;
#pragma unroll
  for (int i = 0; i < MA; ++i) {
    if (i < natm) psum += P[i];
    if (i == own) pown = P[i];
  }
  w[pt] = vol[pt] * pown / psum;
}


// Weight of one point, every atom pair evaluated from scratch (blocks whose
// points straddle two molecules, or molecules above BECKE_MAXA atoms).
template <typename T>
__device__ inline double becke_point(T x, T y, T z, int a0, int na, int own, const int* __restrict__ atm,
                                     const double* __restrict__ env) {
  T psum = T(0), pown = T(0);
  for (int i = a0; i < a0 + na; ++i) {
    const double* Ri = env + atm[i * ATM_SLOTS + 1];
    const T dxi = x - (T)Ri[0], dyi = y - (T)Ri[1], dzi = z - (T)Ri[2];
    const T ri = sqrt(dxi * dxi + dyi * dyi + dzi * dzi);
    const T radi = sqrt((T)c_bragg[atm[i * ATM_SLOTS]]);
    T pi = T(1);
    for (int j = a0; j < a0 + na; ++j) {
      if (j == i) continue;
      const double* Rj = env + atm[j * ATM_SLOTS + 1];
      const T dxj = x - (T)Rj[0], dyj = y - (T)Rj[1], dzj = z - (T)Rj[2];
      const T rj = sqrt(dxj * dxj + dyj * dyj + dzj * dzj);
      const T ex = (T)(Ri[0] - Rj[0]), ey = (T)(Ri[1] - Rj[1]), ez = (T)(Ri[2] - Rj[2]);
      const T radj = sqrt((T)c_bragg[atm[j * ATM_SLOTS]]);
      T aij = T(0.25) * (radj / radi - radi / radj);
      aij = aij < T(-0.5) ? T(-0.5) : (aij > T(0.5) ? T(0.5) : aij);
      T gg = (ri - rj) / sqrt(ex * ex + ey * ey + ez * ez);
      gg = gg + aij * (T(1) - gg * gg);
      pi *= T(0.5) * (T(1) - becke_f3(gg));
    }
    psum += pi;
    if (i == own) pown = pi;
  }
  return (double)(pown / psum);
}

// Becke weights of all points (build_grid.py:34-100).  A workgroup takes 256
// consecutive points; when they belong to one molecule of at most BECKE_MAXA
// atoms (the common case) the pair constants 1/R_ij and a_ij are tabulated in
// LDS once, every point keeps its atom distances in registers, and each
// unordered pair is evaluated once: nu_ji = -nu_ij, so s_ji = 1 - s_ij.
constexpr int BECKE_MAXA = 32;

template <typename T>
__global__ void __launch_bounds__(256) becke_batch_kernel(const BatchMol* __restrict__ mols,
                                                          const int* __restrict__ ptmol,
                                                          const int* __restrict__ owner,
                                                          const double* __restrict__ coords,
                                                          const double* __restrict__ vol, int64_t G,
                                                          const int* __restrict__ atm, const double* __restrict__ env,
                                                          T* __restrict__ w, double* __restrict__ w64) {
  __shared__ T s_inv[BECKE_MAXA * BECKE_MAXA], s_a[BECKE_MAXA * BECKE_MAXA];
  const int64_t blk0 = (int64_t)blockIdx.x * blockDim.x;
  const int64_t last = (blk0 + blockDim.x < G ? blk0 + blockDim.x : G) - 1;
  const int m0 = ptmol[blk0];
  const BatchMol Mo = mols[m0];
  const int a0 = (int)Mo.atm0, na = (int)Mo.natm;
  const bool fast = ptmol[last] == m0 && na <= BECKE_MAXA;  // uniform over the workgroup
  const int64_t pt = blk0 + threadIdx.x;
  if (!fast) {
    if (pt >= G) return;
    const BatchMol Mp = mols[ptmol[pt]];
    const double r = becke_point<T>((T)coords[3 * pt], (T)coords[3 * pt + 1], (T)coords[3 * pt + 2], (int)Mp.atm0,
                                    (int)Mp.natm, owner[pt], atm, env);
    const double wt = vol[pt] * r;
    w[pt] = (T)wt;
    if (w64) w64[pt] = wt;
    return;
  }
  for (int q = threadIdx.x; q < na * na; q += blockDim.x) {
    const int i = q / na, j = q - i * na;
    T inv = T(0), aij = T(0);
    if (i != j) {
      const double* Ri = env + atm[(a0 + i) * ATM_SLOTS + 1];
      const double* Rj = env + atm[(a0 + j) * ATM_SLOTS + 1];
      const T ex = (T)(Ri[0] - Rj[0]), ey = (T)(Ri[1] - Rj[1]), ez = (T)(Ri[2] - Rj[2]);
      inv = T(1) / sqrt(ex * ex + ey * ey + ez * ez);
      const T radi = sqrt((T)c_bragg[atm[(a0 + i) * ATM_SLOTS]]), radj = sqrt((T)c_bragg[atm[(a0 + j) * ATM_SLOTS]]);
      aij = T(0.25) * (radj / radi - radi / radj);
      aij = aij < T(-0.5) ? T(-0.5) : (aij > T(0.5) ? T(0.5) : aij);
    }
    s_inv[i * BECKE_MAXA + j] = inv;
    s_a[i * BECKE_MAXA + j] = aij;
  }
  __syncthreads();
  if (pt >= G) return;
  const T x = (T)coords[3 * pt], y = (T)coords[3 * pt + 1], z = (T)coords[3 * pt + 2];
  T r[BECKE_MAXA], P[BECKE_MAXA];
#pragma unroll
  for (int i = 0; i < BECKE_MAXA; ++i) {
    P[i] = T(1);
    r[i] = T(0);
    if (i < na) {
      const double* Ri = env + atm[(a0 + i) * ATM_SLOTS + 1];
      const T dx = x - (T)Ri[0], dy = y - (T)Ri[1], dz = z - (T)Ri[2];
      r[i] = sqrt(dx * dx + dy * dy + dz * dz);
    }
  }
  // (a uniform branch per i rather than a cut-short loop, so both loops stay
  // unrolled and r, P in registers)
#pragma unroll
  for (int i = 1; i < BECKE_MAXA; ++i) {
    if (i < na) {
#pragma unroll
      for (int j = 0; j < i; ++j) {
        T nu = (r[i] - r[j]) * s_inv[i * BECKE_MAXA + j];
        nu = nu + s_a[i * BECKE_MAXA + j] * (T(1) - nu * nu);
        const T f = becke_f3(nu);
        P[i] *= T(0.5) * (T(1) - f);
        P[j] *= T(0.5) * (T(1) + f);
      }
    }
  }
  const int own = owner[pt] - a0;
  T psum = T(0), pown = T(0);
#pragma unroll
  for (int i = 0; i < BECKE_MAXA; ++i) {
    if (i < na) psum += P[i];
    if (i == own) pown = P[i];
  }
  const double wt = vol[pt] * (double)(pown / psum);
  w[pt] = (T)wt;
  if (w64) w64[pt] = wt;
}

// AO values/gradients of all points; molecule m's block is ao[ao0 + (c*N + mu)*g + p].
template <typename T>
__global__ void __launch_bounds__(256) ao_batch_kernel(const BatchMol* __restrict__ mols, const int* __restrict__ ptmol,
                                                       const double* __restrict__ coords, int64_t G,
                                                       const int* __restrict__ atm, const int* __restrict__ bas,
                                                       const double* __restrict__ env, T* __restrict__ ao) {
  const int64_t pt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= G) return;
  const BatchMol Mo = mols[ptmol[pt]];
  const int64_t g = Mo.ngrid, lp = pt - Mo.grid0;
  const int64_t cs = Mo.nao * g;
  T* base = ao + Mo.ao0;
  const double x = coords[3 * pt], y = coords[3 * pt + 1], z = coords[3 * pt + 2];
  int mu = 0;
  for (int ib = (int)Mo.bas0; ib < (int)(Mo.bas0 + Mo.nbas); ++ib) {
    const int* b = bas + ib * BAS_SLOTS;
    const double* R = env + atm[b[0] * ATM_SLOTS + 1];
    const int l = b[1], np = b[2];
    const double* e = env + b[5];
    const double* c = env + b[6];
    const double dx = x - R[0], dy = y - R[1], dz = z - R[2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    double rad = 0.0, drad = 0.0;
    for (int k = 0; k < np; ++k) {
      const double ek = c[k] * exp(-e[k] * r2);
      rad += ek;
      drad += e[k] * ek;
    }
    drad *= -2.0;
    if (l == 0) {
      rad *= ANG_S;
      drad *= ANG_S;
      base[mu * g + lp] = (T)rad;
      base[cs + mu * g + lp] = (T)(drad * dx);
      base[2 * cs + mu * g + lp] = (T)(drad * dy);
      base[3 * cs + mu * g + lp] = (T)(drad * dz);
      mu += 1;
    } else {
      rad *= ANG_P;
      drad *= ANG_P;
      const double d[3] = {dx, dy, dz};
#pragma unroll
      for (int aa = 0; aa < 3; ++aa) {
        const int64_t o = (int64_t)(mu + aa) * g + lp;
        base[o] = (T)(d[aa] * rad);
#pragma unroll
        for (int q = 0; q < 3; ++q) base[(1 + q) * cs + o] = (T)((aa == q ? rad : 0.0) + d[aa] * d[q] * drad);
      }
      mu += 3;
    }
  }
}

template <typename T, int XT, int NP>
static size_t xc_lds(int N) {
  const size_t main = sizeof(T) * XDReg<T, XT, NP>::lds_elems;
  const size_t tail = sizeof(T) * (size_t)N * N + 8 * sizeof(double);  // V + energy partials after the tiles
  return (main > tail ? main : tail) + 8 * sizeof(double);
}

static int xc_blocks(int g) {
  const int ntiles = (g + XC_TILE - 1) / XC_TILE;
  return ntiles < XC_MAX_BLOCKS ? (ntiles < 1 ? 1 : ntiles) : XC_MAX_BLOCKS;
}

// workspace: nb partial V_xc matrices | nb FP64 energy partials | the chunk
// sums of reduce_parts (each part 256-byte aligned)
struct XcWs {
  size_t e_off, lvl_off, bytes;
};

static XcWs xc_ws(size_t es, int N, int g) {
  const int nb = xc_blocks(g);
  XcWs w;
  w.e_off = red_align(es * (size_t)nb * N * N);
  w.lvl_off = w.e_off + red_align(sizeof(double) * (size_t)nb);
  w.bytes = w.lvl_off + es * (size_t)red_chunks(nb) * N * N;
  return w;
}

template <typename T, int NP>
static int xc_fused_launch(const XCArgs<T>& a, int nb, hipStream_t st) {
  const size_t lds = xc_lds<T, XMTraits<T>::TILE, NP>(a.N);
  PSCF_REQUIRE((int)lds <= xc_max_lds(), "pscf_xc: N=%d needs %zu B LDS", a.N, lds);
  PSCF_SET_LDS((xc_fused_kernel<T, NP>), lds);
  xc_fused_kernel<T, NP><<<nb, XC_THREADS, lds, st>>>(a);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

template <typename T>
static int xc_impl(const T* ao, int64_t sc, int64_t smu, int64_t sp, int g, int N, const T* dm, const T* w,
                   double eps, double* exc, T* vxc, void* ws, size_t ws_bytes, hipStream_t st) {
  PSCF_REQUIRE(N <= XM_MAXN, "pscf_xc: N=%d too large (max %d)", N, XM_MAXN);
  PSCF_REQUIRE(3 * sc + (int64_t)(N - 1) * smu + (int64_t)(g - 1) * sp < ((int64_t)1 << 31),
               "pscf_xc: AO block of %lld elements needs 32-bit offsets", (long long)(4 * sc));
  const int nb = xc_blocks(g);
  const XcWs lay = xc_ws(sizeof(T), N, g);
  PSCF_REQUIRE(ws && ws_bytes >= lay.bytes, "pscf_xc: workspace too small (%zu < %zu)", ws_bytes, lay.bytes);
  XCArgs<T> a;
  a.ao = ao;
  a.sc = sc;
  a.smu = smu;
  a.sp = sp;
  a.g = g;
  a.N = N;
  a.dm = dm;
  a.w = w;
  a.eps = (T)eps;
  a.part_v = (T*)ws;
  a.part_e = (double*)((char*)ws + lay.e_off);
  // FP64: exact atomic sums into the zeroed outputs, no partials to reduce
  constexpr bool ATOMIC = sizeof(T) == 8;
  a.vout = ATOMIC ? vxc : nullptr;
  a.eout = ATOMIC ? exc : nullptr;
  if (ATOMIC) {
    const int rz = zero2((double*)vxc, (int64_t)N * N, exc, 1, st);
    if (rz) return rz;
  }
  int rc;
  switch (xm_np(N)) {
    case 16: rc = xc_fused_launch<T, 16>(a, nb, st); break;
    case 32: rc = xc_fused_launch<T, 32>(a, nb, st); break;
    case 48: rc = xc_fused_launch<T, 48>(a, nb, st); break;
    case 64: rc = xc_fused_launch<T, 64>(a, nb, st); break;
    default: rc = xc_fused_launch<T, 80>(a, nb, st); break;
  }
  if (rc || ATOMIC) return rc;
  return reduce_parts<T, true>(a.part_v, nb, N, (T*)((char*)ws + lay.lvl_off), vxc, a.part_e, exc, st);
}

// ------------------------------------------------------- nuclear gradient ----
// XC part of the nuclear gradient (grad_elec, pyscf_ipu/nanoDFT/nanoDFT.py:
// 403-421: vmat and the second-derivative term, without grid response).  With
// X_i = sum_j D_ij phi_j, the potential weights c = w [v_rho / 2, 4 v_sigma
// grad rho / 2] of the energy path, Y_j = c_0 phi_j + sum_e c_e d_e phi_j and
// YD_i = sum_j D_ij Y_j, each AO i collects
//   t_s[i] = sum_p d_s phi_i YD_i + (c_0 d_s phi_i + sum_e c_e d_s d_e phi_i) X_i
// and the atom A of AO i gets -2 t_s[i].  One workgroup per tile of TP
// points: the AO values, gradients and second derivatives are evaluated in
// the kernel (LDS [10][TP][N]), X, Y and YD live in LDS, D in LDS when it
// fits; per-tile partials t[tile][N][3] are summed in tile order by
// xc_grad_reduce_kernel.
constexpr int XCG_THREADS = 256;

// second-derivative component of (s, e) after value and gradient: xx xy xz yy yz zz
__device__ inline int d2_comp(int s, int e) {
  const int a = s < e ? s : e, b = s < e ? e : s;
  return 4 + (a == 0 ? b : (a == 1 ? 2 + b : 5));
}

__global__ void __launch_bounds__(XCG_THREADS) xc_grad_kernel(const double* __restrict__ coords,
                                                              const double* __restrict__ weights, int g,
                                                              const int* __restrict__ atm,
                                                              const int* __restrict__ bas, int nbas,
                                                              const double* __restrict__ env, int N,
                                                              const double* __restrict__ dm_g, int dm_lds, int TP,
                                                              double eps, double* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int TN = TP * N;
  double* Phi = reinterpret_cast<double*>(smem);  // [10][TP][N]
  double* X = Phi + 10 * TN;                       // [TP][N]
  double* Y = X + TN;                              // [TP][N]
  double* YD = Y + TN;                             // [TP][N]
  double* coef = YD + TN;                          // [4][TP]
  double* Dl = coef + 4 * TP;                      // [N][N] when dm_lds
  int* ao_loc = reinterpret_cast<int*>(Dl + (dm_lds ? N * N : 0));
  const double* D = dm_lds ? Dl : dm_g;
  const int tid = threadIdx.x;
  if (dm_lds)
    for (int x = tid; x < N * N; x += blockDim.x) Dl[x] = dm_g[x];
  if (tid == 0) {
    int acc = 0;
    for (int i = 0; i < nbas; ++i) {
      ao_loc[i] = acc;
      acc += 2 * bas[i * BAS_SLOTS + 1] + 1;
    }
    ao_loc[nbas] = acc;
  }
  __syncthreads();
  const int p0 = blockIdx.x * TP;
  // AO values, gradients and second derivatives: item (point, shell)
  for (int it = tid; it < TP * nbas; it += blockDim.x) {
    const int p = it / nbas, ib = it - p * nbas;
    const int pt = p0 + p;
    const int* b = bas + ib * BAS_SLOTS;
    const int mu = ao_loc[ib];
    const int l = b[1];
    double f0 = 0.0, f1 = 0.0, f2 = 0.0, dd[3] = {0.0, 0.0, 0.0};
    if (pt < g) {
      const double* R = env + atm[b[0] * ATM_SLOTS + 1];
      const double* e = env + b[5];
      const double* c = env + b[6];
      for (int d = 0; d < 3; ++d) dd[d] = coords[3 * pt + d] - R[d];
      const double r2 = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
      for (int k = 0; k < b[2]; ++k) {
        const double ek = c[k] * exp(-e[k] * r2);
        f0 += ek;
        f1 += e[k] * ek;
        f2 += e[k] * e[k] * ek;
      }
      const double ang = l ? ANG_P : ANG_S;
      f0 *= ang;
      f1 *= -2.0 * ang;
      f2 *= 4.0 * ang;
    }
    if (l == 0) {
      Phi[p * N + mu] = f0;
#pragma unroll
      for (int x = 0; x < 3; ++x) Phi[(1 + x) * TN + p * N + mu] = f1 * dd[x];
#pragma unroll
      for (int x = 0; x < 3; ++x)
#pragma unroll
        for (int y = x; y < 3; ++y) Phi[d2_comp(x, y) * TN + p * N + mu] = f2 * dd[x] * dd[y] + (x == y ? f1 : 0.0);
    } else {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int o = p * N + mu + a;
        Phi[o] = dd[a] * f0;
#pragma unroll
        for (int x = 0; x < 3; ++x) Phi[(1 + x) * TN + o] = (a == x ? f0 : 0.0) + dd[a] * f1 * dd[x];
#pragma unroll
        for (int x = 0; x < 3; ++x)
#pragma unroll
          for (int y = x; y < 3; ++y) {
            double v = dd[a] * (f2 * dd[x] * dd[y] + (x == y ? f1 : 0.0));
            if (a == x) v += f1 * dd[y];
            if (a == y) v += f1 * dd[x];
            Phi[d2_comp(x, y) * TN + o] = v;
          }
      }
    }
  }
  __syncthreads();
  // X = phi0 D
  for (int it = tid; it < TN; it += blockDim.x) {
    const int p = it / N, j = it - p * N;
    double v = 0.0;
    for (int k = 0; k < N; ++k) v += Phi[p * N + k] * D[k * N + j];
    X[it] = v;
  }
  __syncthreads();
  // rho_c = sum_j X phi_c, B3LYP, potential weights
  if (tid < TP) {
    const int pt = p0 + tid;
    double rc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double v = 0.0;
      for (int j = 0; j < N; ++j) v += X[tid * N + j] * Phi[c * TN + tid * N + j];
      rc[c] = v;
    }
    double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
    if (pt < g) {
      double e, vr, vn;
      b3lyp_point<double>(rc[0], rc[1], rc[2], rc[3], eps, e, vr, vn);
      const double wt = weights[pt];
      c0 = 0.5 * vr * wt;
      const double f = 4.0 * vn * wt;
      c1 = f * rc[1];
      c2 = f * rc[2];
      c3 = f * rc[3];
    }
    coef[tid] = c0;
    coef[TP + tid] = c1;
    coef[2 * TP + tid] = c2;
    coef[3 * TP + tid] = c3;
  }
  __syncthreads();
  for (int it = tid; it < TN; it += blockDim.x) {
    const int p = it / N;
    Y[it] = coef[p] * Phi[it] + coef[TP + p] * Phi[TN + it] + coef[2 * TP + p] * Phi[2 * TN + it] +
            coef[3 * TP + p] * Phi[3 * TN + it];
  }
  __syncthreads();
  for (int it = tid; it < TN; it += blockDim.x) {
    const int p = it / N, j = it - p * N;
    double v = 0.0;
    for (int k = 0; k < N; ++k) v += Y[p * N + k] * D[k * N + j];
    YD[it] = v;
  }
  __syncthreads();
  // per AO, summed over the tile's points in order
  for (int i = tid; i < N; i += blockDim.x) {
    double t[3] = {0.0, 0.0, 0.0};
    for (int p = 0; p < TP; ++p) {
      const int o = p * N + i;
      const double x = X[o], yd = YD[o], c0 = coef[p];
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        double z = c0 * Phi[(1 + s) * TN + o];
#pragma unroll
        for (int e = 0; e < 3; ++e) z += coef[(1 + e) * TP + p] * Phi[d2_comp(s, e) * TN + o];
        t[s] += Phi[(1 + s) * TN + o] * yd + z * x;
      }
    }
    double* dst = part + ((int64_t)blockIdx.x * N + i) * 3;
    dst[0] = t[0];
    dst[1] = t[1];
    dst[2] = t[2];
  }
}

// grad[atom][s] += -2 sum_{AO i of atom} sum_tiles part[tile][i][s] (fixed order)
__global__ void __launch_bounds__(256) xc_grad_reduce_kernel(const double* __restrict__ part, int ntiles, int N,
                                                             const int* __restrict__ bas, int nbas, int natm,
                                                             double* __restrict__ grad) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* tot = reinterpret_cast<double*>(smem);  // [N][3]
  for (int x = threadIdx.x; x < 3 * N; x += blockDim.x) {
    double v = 0.0;
    for (int t = 0; t < ntiles; ++t) v += part[(int64_t)t * 3 * N + x];
    tot[x] = v;
  }
  __syncthreads();
  for (int x = threadIdx.x; x < 3 * natm; x += blockDim.x) {
    const int a = x / 3, s = x - 3 * a;
    double v = 0.0;
    int mu = 0;
    for (int ib = 0; ib < nbas; ++ib) {
      const int w = 2 * bas[ib * BAS_SLOTS + 1] + 1;
      if (bas[ib * BAS_SLOTS] == a)
        for (int k = 0; k < w; ++k) v += tot[(mu + k) * 3 + s];
      mu += w;
    }
    grad[x] += -2.0 * v;
  }
}

static size_t xc_grad_lds(int N, int nbas, int TP, bool dm_lds) {
  return sizeof(double) * (13 * (size_t)TP * N + 4 * (size_t)TP + (dm_lds ? (size_t)N * N : 0)) +
         sizeof(int) * (nbas + 1);
}

// Largest tile (points) whose LDS fits, D in LDS if that still allows 8 points.
static void xc_grad_shape(int N, int nbas, int& TP, bool& dm_lds) {
  const size_t cap = (size_t)xc_max_lds();
  dm_lds = xc_grad_lds(N, nbas, 8, true) <= cap;
  TP = 0;
  for (int t = 32; t >= 1; t >>= 1)
    if (xc_grad_lds(N, nbas, t, dm_lds) <= cap) {
      TP = t;
      break;
    }
}

}  // namespace pscf

using namespace pscf;

extern "C" int pscf_ao_eval(int dtype, const int* d_atm, int natm, const int* d_bas, int nbas, const double* d_env,
                            int nao, const double* d_coords, int g, void* d_ao, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_ao_eval: bad dtype");
  PSCF_REQUIRE(natm > 0 && nbas > 0 && nao > 0 && g >= 0, "pscf_ao_eval: bad sizes");
  if (g == 0) return PSCF_OK;
  PSCF_REQUIRE(d_atm && d_bas && d_env && d_coords && d_ao, "pscf_ao_eval: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64)
    ao_eval_kernel<double><<<(g + 255) / 256, 256, 0, st>>>(d_atm, d_bas, nbas, d_env, d_coords, g, nao, (double*)d_ao);
  else
    ao_eval_kernel<float><<<(g + 255) / 256, 256, 0, st>>>(d_atm, d_bas, nbas, d_env, d_coords, g, nao, (float*)d_ao);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_becke_weights(const double* d_coords, const int* d_owner, const double* d_vol, int g,
                                  const double* d_atom_xyz, const double* d_adjust, int natm, double* d_w,
                                  void* stream) {
  PSCF_REQUIRE(g >= 0 && natm > 0 && natm <= 88, "pscf_becke_weights: natm=%d out of range (1..88)", natm);
  if (g == 0) return PSCF_OK;
  PSCF_REQUIRE(d_coords && d_owner && d_vol && d_atom_xyz && d_adjust && d_w, "pscf_becke_weights: null pointer");
  if (natm <= BECKE_FAST_MAXA) {
    becke_fast_kernel<<<(g + 255) / 256, 256, 0, (hipStream_t)stream>>>(d_coords, d_owner, d_vol, g, d_atom_xyz,
                                                                       d_adjust, natm, d_w);
    PSCF_HIP(hipGetLastError());
    return PSCF_OK;
  }
  const size_t lds = sizeof(double) * (3 * (size_t)natm + (size_t)natm * natm);
  becke_kernel<<<(g + 255) / 256, 256, lds, (hipStream_t)stream>>>(d_coords, d_owner, d_vol, g, d_atom_xyz, d_adjust,
                                                                  natm, d_w);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_b3lyp(int dtype, const void* d_rho, int g, double eps, void* d_exc, void* d_vrho, void* d_vnorm,
                          void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_b3lyp: bad dtype");
  PSCF_REQUIRE(g >= 0, "pscf_b3lyp: g < 0");
  if (g == 0) return PSCF_OK;
  PSCF_REQUIRE(d_rho && d_exc && d_vrho && d_vnorm, "pscf_b3lyp: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64)
    b3lyp_kernel<double><<<(g + 255) / 256, 256, 0, st>>>((const double*)d_rho, g, eps, (double*)d_exc,
                                                          (double*)d_vrho, (double*)d_vnorm);
  else
    b3lyp_kernel<float><<<(g + 255) / 256, 256, 0, st>>>((const float*)d_rho, g, (float)eps, (float*)d_exc,
                                                         (float*)d_vrho, (float*)d_vnorm);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

// Host evaluation of the same pointwise code (FP64), for CPU-side checks of
// the hand-derived derivatives.
extern "C" int pscf_b3lyp_host(const double* h_rho, int g, double eps, double* h_exc, double* h_vrho,
                               double* h_vnorm) {
  PSCF_REQUIRE(g >= 0 && (g == 0 || (h_rho && h_exc && h_vrho && h_vnorm)), "pscf_b3lyp_host: bad arguments");
  for (int p = 0; p < g; ++p)
    b3lyp_point<double>(h_rho[p], h_rho[g + p], h_rho[2 * g + p], h_rho[3 * g + p], eps, h_exc[p], h_vrho[p],
                        h_vnorm[p]);
  return PSCF_OK;
}

extern "C" size_t pscf_xc_workspace_bytes(int dtype, int N, int g) {
  return xc_ws(dtype == PSCF_F64 ? 8 : 4, N, g).bytes;
}

extern "C" int pscf_xc(int dtype, const void* d_ao, int64_t ao_sc, int64_t ao_smu, int64_t ao_sp, int g, int N,
                       const void* d_dm, const void* d_weights, double eps, double* d_exc, void* d_vxc, void* d_ws,
                       size_t ws_bytes, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_xc: bad dtype");
  PSCF_REQUIRE(g > 0 && N > 0, "pscf_xc: empty grid or basis");
  PSCF_REQUIRE(d_ao && d_dm && d_weights && d_exc && d_vxc, "pscf_xc: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64)
    return xc_impl<double>((const double*)d_ao, ao_sc, ao_smu, ao_sp, g, N, (const double*)d_dm,
                           (const double*)d_weights, eps, d_exc, (double*)d_vxc, d_ws, ws_bytes, st);
  return xc_impl<float>((const float*)d_ao, ao_sc, ao_smu, ao_sp, g, N, (const float*)d_dm, (const float*)d_weights,
                        eps, d_exc, (float*)d_vxc, d_ws, ws_bytes, st);
}

extern "C" int pscf_batch_grid(const int* d_work, int nwork, const int* d_atom_elem, const int64_t* d_atom_pt0,
                               const int* d_atom_mol, const int* d_tmpl_off, const int* d_tmpl_n, const double* d_tmpl,
                               const int* d_atm, const double* d_env, double* d_coords, double* d_vol, int* d_owner,
                               int* d_ptmol, void* stream) {
  PSCF_REQUIRE(nwork >= 0, "pscf_batch_grid: nwork < 0");
  if (nwork == 0) return PSCF_OK;
  grid_expand_kernel<<<nwork, 256, 0, (hipStream_t)stream>>>(d_work, d_atom_elem, d_atom_pt0, d_atom_mol, d_tmpl_off,
                                                            d_tmpl_n, d_tmpl, d_atm, d_env, d_coords, d_vol, d_owner,
                                                            d_ptmol);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_batch_becke(int dtype, const void* d_mols, const int* d_ptmol, const int* d_owner,
                                const double* d_coords, const double* d_vol, int64_t G, const int* d_atm,
                                const double* d_env, void* d_w, double* d_w64, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_batch_becke: bad dtype");
  PSCF_REQUIRE(G >= 0, "pscf_batch_becke: G < 0");
  if (G == 0) return PSCF_OK;
  const int64_t blocks = (G + 255) / 256;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64)
    becke_batch_kernel<double><<<(unsigned)blocks, 256, 0, st>>>((const BatchMol*)d_mols, d_ptmol, d_owner, d_coords,
                                                                 d_vol, G, d_atm, d_env, (double*)d_w, d_w64);
  else
    becke_batch_kernel<float><<<(unsigned)blocks, 256, 0, st>>>((const BatchMol*)d_mols, d_ptmol, d_owner, d_coords,
                                                                d_vol, G, d_atm, d_env, (float*)d_w, d_w64);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_batch_ao(int dtype, const void* d_mols, const int* d_ptmol, const double* d_coords, int64_t G,
                             const int* d_atm, const int* d_bas, const double* d_env, void* d_ao, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_batch_ao: bad dtype");
  PSCF_REQUIRE(G >= 0, "pscf_batch_ao: G < 0");
  if (G == 0) return PSCF_OK;
  const int64_t blocks = (G + 255) / 256;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64)
    ao_batch_kernel<double><<<(unsigned)blocks, 256, 0, st>>>((const BatchMol*)d_mols, d_ptmol, d_coords, G, d_atm,
                                                              d_bas, d_env, (double*)d_ao);
  else
    ao_batch_kernel<float><<<(unsigned)blocks, 256, 0, st>>>((const BatchMol*)d_mols, d_ptmol, d_coords, G, d_atm,
                                                             d_bas, d_env, (float*)d_ao);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

// LDS of the batched kernel: the tile pipeline (or, after the tiles, V and
// the energy partials), then in the direct form the shell table.
template <typename T, int XT, int NP>
static size_t xc_batch_lds(int max_nao, const XcDirect* da) {
  const size_t base = (xc_lds<T, XT, NP>(max_nao) + 31) & ~(size_t)31;
  return da ? base + xc_shell_bytes(da->nsh_max, da->kp, sizeof(T)) : base;
}

template <typename T, int XT, int NP>
static int batch_xc_launch(const void* d_mols, const int* d_order, int nmol, int max_nao, const double* d_ctl,
                           const void* d_ao, const XcDirect* da, const void* d_dm, const void* d_w, double eps,
                           void* d_vxc, double* d_exc, int split, void* d_part_v, double* d_part_e,
                           int64_t part_v_elems, int64_t part_e_elems, hipStream_t st) {
  const size_t lds = xc_batch_lds<T, XT, NP>(max_nao, da);
  PSCF_REQUIRE((int)lds <= xc_max_lds(), "pscf_batch_xc: nao %d needs %zu B LDS", max_nao, lds);
  const dim3 grid(nmol, split);
  if (da) {
    XcDirect a = *da;
    a.tab_off = (int)((xc_lds<T, XT, NP>(max_nao) + 31) & ~(size_t)31);
    PSCF_SET_LDS((xc_batch_kernel<T, XT, NP, true>), lds);
    xc_batch_kernel<T, XT, NP, true><<<grid, XC_THREADS, lds, st>>>(
        (const BatchMol*)d_mols, d_order, max_nao, d_ctl, nullptr, a, (const T*)d_dm, (const T*)d_w, (T)eps,
        (T*)d_vxc, d_exc, (T*)d_part_v, d_part_e, part_v_elems, part_e_elems);
  } else {
    PSCF_SET_LDS((xc_batch_kernel<T, XT, NP, false>), lds);
    xc_batch_kernel<T, XT, NP, false><<<grid, XC_THREADS, lds, st>>>(
        (const BatchMol*)d_mols, d_order, max_nao, d_ctl, (const T*)d_ao, XcDirect{}, (const T*)d_dm,
        (const T*)d_w, (T)eps, (T*)d_vxc, d_exc, (T*)d_part_v, d_part_e, part_v_elems, part_e_elems);
  }
  PSCF_HIP(hipGetLastError());
  xc_batch_reduce_kernel<T><<<nmol, 256, sizeof(T) * (size_t)max_nao * max_nao, st>>>(
      (const BatchMol*)d_mols, d_order, max_nao, d_ctl, (const T*)d_part_v, d_part_e, part_v_elems, part_e_elems,
      (T*)d_vxc, d_exc, da ? da->nsh_max : 0);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

template <typename T, int XT>
static int batch_xc_np(int np, const void* d_mols, const int* d_order, int nmol, int max_nao, const double* d_ctl,
                       const void* d_ao, const XcDirect* da, const void* d_dm, const void* d_w, double eps,
                       void* d_vxc, double* d_exc, int split, void* d_part_v, double* d_part_e, int64_t pv,
                       int64_t pe, hipStream_t st) {
#define PSCF_XC_NP(V)                                                                                           \
  case V:                                                                                                       \
    return batch_xc_launch<T, XT, V>(d_mols, d_order, nmol, max_nao, d_ctl, d_ao, da, d_dm, d_w, eps, d_vxc,  \
                                     d_exc, split, d_part_v, d_part_e, pv, pe, st);
  switch (np) {
    PSCF_XC_NP(16)
    PSCF_XC_NP(32)
    PSCF_XC_NP(48)
    PSCF_XC_NP(64)
    PSCF_XC_NP(80)
  }
#undef PSCF_XC_NP
  PSCF_REQUIRE(false, "pscf_batch_xc: nao padding %d unsupported", np);
  return PSCF_EINVAL;
}

static int batch_xc_any(int dtype, const void* d_mols, const int* d_order, int nmol, int max_nao,
                        const double* d_ctl, const void* d_ao, const XcDirect* da, const void* d_dm, const void* d_w,
                        double eps, void* d_vxc, double* d_exc, int split, int tile, void* d_part_v,
                        double* d_part_e, int64_t part_v_elems, int64_t part_e_elems, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_batch_xc: bad dtype");
  PSCF_REQUIRE(nmol > 0 && max_nao > 0 && d_mols && d_order && (d_ao || da) && d_dm && d_w && d_vxc && d_exc,
               "pscf_batch_xc: bad arguments");
  PSCF_REQUIRE(split >= 1 && split <= 64, "pscf_batch_xc: bad split %d", split);
  PSCF_REQUIRE(d_part_v && d_part_e && part_v_elems >= 0 && part_e_elems >= 0, "pscf_batch_xc: no chunk buffers");
  PSCF_REQUIRE(tile == 32 || (tile == 64 && dtype == PSCF_F32), "pscf_batch_xc: tile %d (32, or 64 in FP32)", tile);
  PSCF_REQUIRE(max_nao <= XM_MAXN, "pscf_batch_xc: nao %d too large", max_nao);
  hipStream_t st = (hipStream_t)stream;
  const int np = xm_np(max_nao);
  if (dtype == PSCF_F64)
    return batch_xc_np<double, 32>(np, d_mols, d_order, nmol, max_nao, d_ctl, d_ao, da, d_dm, d_w, eps, d_vxc,
                                   d_exc, split, d_part_v, d_part_e, part_v_elems, part_e_elems, st);
  if (tile == 64)
    return batch_xc_np<float, 64>(np, d_mols, d_order, nmol, max_nao, d_ctl, d_ao, da, d_dm, d_w, eps, d_vxc,
                                  d_exc, split, d_part_v, d_part_e, part_v_elems, part_e_elems, st);
  return batch_xc_np<float, 32>(np, d_mols, d_order, nmol, max_nao, d_ctl, d_ao, da, d_dm, d_w, eps, d_vxc, d_exc,
                                split, d_part_v, d_part_e, part_v_elems, part_e_elems, st);
}

extern "C" int pscf_batch_xc(int dtype, const void* d_mols, const int* d_order, int nmol, int max_nao,
                             const double* d_ctl, const void* d_ao, const void* d_dm, const void* d_w, double eps,
                             void* d_vxc, double* d_exc, int split, int tile, void* d_part_v, double* d_part_e,
                             int64_t part_v_elems, int64_t part_e_elems, void* stream) {
  PSCF_REQUIRE(d_ao, "pscf_batch_xc: no AO block");
  return batch_xc_any(dtype, d_mols, d_order, nmol, max_nao, d_ctl, d_ao, nullptr, d_dm, d_w, eps, d_vxc, d_exc,
                      split, tile, d_part_v, d_part_e, part_v_elems, part_e_elems, stream);
}

extern "C" int pscf_batch_xc_direct(int dtype, const void* d_mols, const int* d_order, int nmol, int max_nao,
                                    const double* d_ctl, const double* d_coords, const int* d_atm, const int* d_bas,
                                    const double* d_env, int nsh_max, int kp, const void* d_dm, const void* d_w,
                                    double eps, void* d_vxc, double* d_exc, int split, int tile, void* d_part_v,
                                    double* d_part_e, int64_t part_v_elems, int64_t part_e_elems, void* stream) {
  PSCF_REQUIRE(d_coords && d_atm && d_bas && d_env, "pscf_batch_xc_direct: null pointer");
  PSCF_REQUIRE(nsh_max >= 1 && nsh_max <= XM_MAXN && kp >= 1 && kp <= 16,
               "pscf_batch_xc_direct: nsh_max %d (1..%d) or kp %d (1..16) out of range", nsh_max, XM_MAXN, kp);
  // at least three primitive slots per record: the stage reads the first three unconditionally
  XcDirect da{d_coords, d_atm, d_bas, d_env, nsh_max, kp < 3 ? 3 : kp, 0};
  return batch_xc_any(dtype, d_mols, d_order, nmol, max_nao, d_ctl, nullptr, &da, d_dm, d_w, eps, d_vxc, d_exc,
                      split, tile, d_part_v, d_part_e, part_v_elems, part_e_elems, stream);
}

extern "C" size_t pscf_grad_xc_workspace_bytes(int nao, int nbas, int g) {
  int TP;
  bool dm_lds;
  xc_grad_shape(nao, nbas, TP, dm_lds);
  if (TP == 0) return 0;
  return sizeof(double) * 3 * (size_t)nao * ((g + TP - 1) / TP);
}

extern "C" int pscf_grad_xc(const double* d_coords, const double* d_weights, int g, const int* d_atm, int natm,
                            const int* d_bas, int nbas, const double* d_env, int nao, const double* d_dm, double eps,
                            void* d_ws, size_t ws_bytes, double* d_grad, void* stream) {
  PSCF_REQUIRE(g > 0 && natm > 0 && nbas > 0 && nao > 0 && nao <= 4096, "pscf_grad_xc: bad sizes");
  PSCF_REQUIRE(d_coords && d_weights && d_atm && d_bas && d_env && d_dm && d_ws && d_grad, "pscf_grad_xc: null pointer");
  int TP;
  bool dm_lds;
  xc_grad_shape(nao, nbas, TP, dm_lds);
  PSCF_REQUIRE(TP > 0, "pscf_grad_xc: nao %d does not fit the LDS", nao);
  const int ntiles = (g + TP - 1) / TP;
  PSCF_REQUIRE(ws_bytes >= sizeof(double) * 3 * (size_t)nao * ntiles, "pscf_grad_xc: workspace too small");
  PSCF_REQUIRE((size_t)sizeof(double) * 3 * nao <= (size_t)xc_max_lds(), "pscf_grad_xc: nao %d too large", nao);
  hipStream_t st = (hipStream_t)stream;
  const size_t lds = xc_grad_lds(nao, nbas, TP, dm_lds);
  PSCF_SET_LDS(xc_grad_kernel, lds);
  xc_grad_kernel<<<ntiles, XCG_THREADS, lds, st>>>(d_coords, d_weights, g, d_atm, d_bas, nbas, d_env, nao, d_dm,
                                                   dm_lds ? 1 : 0, TP, eps, (double*)d_ws);
  PSCF_HIP(hipGetLastError());
  const size_t lds2 = sizeof(double) * 3 * (size_t)nao;
  PSCF_SET_LDS(xc_grad_reduce_kernel, lds2);
  xc_grad_reduce_kernel<<<1, 256, lds2, st>>>((const double*)d_ws, ntiles, nao, d_bas, nbas, natm, d_grad);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}
