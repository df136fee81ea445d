// Kernels for molecules too large for the LDS-resident kernels (6-31G, nao
// 130-200), used by the precision study:
//   * packed s8 ERIs -> the full s1 tensor (pq|rs) = T[p][q][r][s];
//   * J/K from the s1 tensor in one streaming pass (a workgroup per row p:
//     J[p][q] = <T[p][q], D>, K[p][r] = sum_q T[p][q][r][:] . D[q][:]);
//   * the grid XC pointwise stages around the AO-grid GEMMs (pscf_gemm):
//     rho_c[p] = sum_nu X[nu][p] phi_c[nu][p] with X = D phi0, B3LYP, and
//     Y[nu][p] = sum_c coef_c[p] phi_c[nu][p] for V = phi0 Y^T;
//   * casts to the GEMM input precision (BF16 / FP16 / FP32 / FP64);
//   * the SCF step with its working matrices in a global workspace (L2).
#include <type_traits>

#include "b3lyp.h"
#include "common.h"

namespace pscf {

enum LType : int { L_BF16 = 0, L_F16 = 1, L_F32 = 2, L_F64 = 3 };

__device__ inline unsigned short l_bf16(double x) {  // nearest even, via FP32
  unsigned u = __float_as_uint((float)x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

template <int TY>
struct LT;
template <>
struct LT<L_BF16> {
  typedef unsigned short t;
  static __device__ inline t to(double x) { return l_bf16(x); }
  static __device__ inline double from(t v) { return (double)__uint_as_float(((unsigned)v) << 16); }
};
template <>
struct LT<L_F16> {
  typedef _Float16 t;
  static __device__ inline t to(double x) { return (_Float16)(float)x; }
  static __device__ inline double from(t v) { return (double)(float)v; }
};
template <>
struct LT<L_F32> {
  typedef float t;
  static __device__ inline t to(double x) { return (float)x; }
  static __device__ inline double from(t v) { return (double)v; }
};
template <>
struct LT<L_F64> {
  typedef double t;
  static __device__ inline t to(double x) { return x; }
  static __device__ inline double from(t v) { return v; }
};

template <int TY>
__global__ void cast_kernel(int64_t n, const double* __restrict__ x, typename LT<TY>::t* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = LT<TY>::to(x[i]);
}

__device__ inline int64_t tri_root64(int64_t c) {
  int64_t a = (int64_t)((sqrt(8.0 * (double)c + 1.0) - 1.0) * 0.5);
  while ((a + 1) * (a + 2) / 2 <= c) ++a;
  while (a * (a + 1) / 2 > c) --a;
  return a;
}

template <typename TS, typename TD>
__global__ void s8_to_s1_kernel(int N, int64_t n_s8, const TS* __restrict__ s8, TD* __restrict__ s1) {
  const int64_t NN = (int64_t)N * N;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n_s8; x += (int64_t)gridDim.x * blockDim.x) {
    const int64_t pq = tri_root64(x), rs = x - pq * (pq + 1) / 2;
    const int64_t p = tri_root64(pq), q = pq - p * (p + 1) / 2;
    const int64_t r = tri_root64(rs), s = rs - r * (r + 1) / 2;
    const TD v = (TD)s8[x];
    const int64_t ab[2] = {p * N + q, q * N + p}, cd[2] = {r * N + s, s * N + r};
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        s1[ab[i] * NN + cd[j]] = v;
        s1[cd[j] * NN + ab[i]] = v;
      }
  }
}

__device__ inline double l_wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// dJK[p][:] = J[p][:] - hhalf K[p][:], one workgroup per row p.
template <typename T>
__global__ void __launch_bounds__(256) jk_s1_kernel(int N, const T* __restrict__ s1, const double* __restrict__ D,
                                                    double hhalf, double* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* kacc = reinterpret_cast<double*>(smem);
  double* jacc = kacc + N;
  double* dq = jacc + N;
  const int p = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = blockDim.x >> 6;
  for (int i = tid; i < N; i += blockDim.x) {
    kacc[i] = 0.0;
    jacc[i] = 0.0;
  }
  for (int q = 0; q < N; ++q) {
    __syncthreads();
    for (int s = tid; s < N; s += blockDim.x) dq[s] = D[(int64_t)q * N + s];
    __syncthreads();
    const T* blk = s1 + ((int64_t)p * N + q) * (int64_t)N * N;
    double jpart = 0.0;
    for (int r = wave; r < N; r += nwave) {
      const T* row = blk + (int64_t)r * N;
      const double* drow = D + (int64_t)r * N;
      double kp = 0.0;
      for (int s = lane; s < N; s += 64) {
        const double v = (double)row[s];
        kp += v * dq[s];
        jpart += v * drow[s];
      }
      kp = l_wave_sum(kp);
      if (lane == 0) kacc[r] += kp;  // row r belongs to this wavefront alone
    }
    jpart = l_wave_sum(jpart);
    if (lane == 0) atomicAdd(jacc + q, jpart);
  }
  __syncthreads();
  for (int r = tid; r < N; r += blockDim.x) out[(int64_t)p * N + r] = jacc[r] - hhalf * kacc[r];
}

// Per point p: rho_c = sum_nu X[nu][p] phi_c[nu][p] (c = 0: rho, 1..3: grad/2),
// B3LYP, E_xc partial sums per block, and the V_xc coefficients
// coef = (v_rho w / 2, 4 v_sigma w rho_x, .., ..).  phi: [4][N][g] in type TY.
template <int TY>
__global__ void __launch_bounds__(256) xc_rho_large_kernel(int N, int g, const double* __restrict__ X,
                                                           const typename LT<TY>::t* __restrict__ phi,
                                                           const double* __restrict__ w, double eps,
                                                           double* __restrict__ coef, double* __restrict__ e_part) {
  __shared__ double red[4];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double el = 0.0;
  if (p < g) {
    const int64_t sc = (int64_t)N * g;
    double r[4] = {0.0, 0.0, 0.0, 0.0};
    for (int nu = 0; nu < N; ++nu) {
      const int64_t o = (int64_t)nu * g + p;
      const double x = X[o];
#pragma unroll
      for (int c = 0; c < 4; ++c) r[c] += x * LT<TY>::from(phi[c * sc + o]);
    }
    double e, vr, vn;
    b3lyp_point<double>(r[0], r[1], r[2], r[3], eps, e, vr, vn);
    const double wt = w[p];
    el = r[0] * wt * e;
    const double f = 4.0 * vn * wt;
    coef[p] = 0.5 * vr * wt;
    coef[(int64_t)g + p] = f * r[1];
    coef[2 * (int64_t)g + p] = f * r[2];
    coef[3 * (int64_t)g + p] = f * r[3];
  }
  el = l_wave_sum(el);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = el;
  __syncthreads();
  if (threadIdx.x == 0) e_part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

template <int TY>
__global__ void __launch_bounds__(256) xc_y_large_kernel(int N, int g, const double* __restrict__ coef,
                                                         const typename LT<TY>::t* __restrict__ phi,
                                                         typename LT<TY>::t* __restrict__ Y) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g) return;
  const int64_t sc = (int64_t)N * g;
  const double c0 = coef[p], c1 = coef[(int64_t)g + p], c2 = coef[2 * (int64_t)g + p], c3 = coef[3 * (int64_t)g + p];
  for (int nu = 0; nu < N; ++nu) {
    const int64_t o = (int64_t)nu * g + p;
    const double y = c0 * LT<TY>::from(phi[o]) + c1 * LT<TY>::from(phi[sc + o]) + c2 * LT<TY>::from(phi[2 * sc + o]) +
                     c3 * LT<TY>::from(phi[3 * sc + o]);
    Y[o] = LT<TY>::to(y);
  }
}

// E_xc = sum of the block partials (Neumaier) and V_xc = Vh + Vh^T.
__global__ void __launch_bounds__(256) xc_finish_large_kernel(int N, const double* __restrict__ Vh,
                                                              const double* __restrict__ e_part, int nparts,
                                                              double* __restrict__ vxc, double* __restrict__ exc) {
  for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < N * N; x += gridDim.x * blockDim.x) {
    const int i = x / N, j = x - i * N;
    vxc[x] = Vh[x] + Vh[j * N + i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double s = 0.0, c = 0.0;
    for (int k = 0; k < nparts; ++k) {
      const double a = e_part[k], t = s + a;
      c += fabs(s) >= fabs(a) ? (s - t) + a : (a - t) + s;
      s = t;
    }
    *exc = s + c;
  }
}

template <int TY>
static int cast_launch(int64_t n, const double* x, void* y, hipStream_t st) {
  const int64_t blocks = (n + 255) / 256;
  cast_kernel<TY><<<(unsigned)(blocks < 65536 ? blocks : 65536), 256, 0, st>>>(n, x, (typename LT<TY>::t*)y);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

template <int TY>
static int xc_large_launch(int stage, int N, int g, const double* X, const void* phi, const double* w, double eps,
                           double* coef, double* e_part, void* Y, hipStream_t st) {
  typedef typename LT<TY>::t T;
  const unsigned blocks = (unsigned)((g + 255) / 256);
  if (stage == 0)
    xc_rho_large_kernel<TY><<<blocks, 256, 0, st>>>(N, g, X, (const T*)phi, w, eps, coef, e_part);
  else
    xc_y_large_kernel<TY><<<blocks, 256, 0, st>>>(N, g, coef, (const T*)phi, (T*)Y);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

}  // namespace pscf

using namespace pscf;

extern "C" int pscf_cast_to(int out_type, const double* d_x, void* d_y, int64_t n, void* stream) {
  PSCF_REQUIRE(n >= 0 && (n == 0 || (d_x && d_y)), "pscf_cast_to: bad arguments");
  if (n == 0) return PSCF_OK;
  hipStream_t st = (hipStream_t)stream;
  switch (out_type) {
    case L_BF16: return cast_launch<L_BF16>(n, d_x, d_y, st);
    case L_F16: return cast_launch<L_F16>(n, d_x, d_y, st);
    case L_F32: return cast_launch<L_F32>(n, d_x, d_y, st);
    case L_F64: return cast_launch<L_F64>(n, d_x, d_y, st);
    default: break;
  }
  PSCF_REQUIRE(false, "pscf_cast_to: bad type %d", out_type);
  return PSCF_EINVAL;
}

extern "C" int pscf_eri_s8_to_s1(int dtype_in, const void* d_s8, int dtype_out, void* d_s1, int N, void* stream) {
  PSCF_REQUIRE((dtype_in == PSCF_F64 || dtype_in == PSCF_F32) && (dtype_out == PSCF_F64 || dtype_out == PSCF_F32),
               "pscf_eri_s8_to_s1: bad dtype");
  PSCF_REQUIRE(N > 0 && N <= 4096 && d_s8 && d_s1, "pscf_eri_s8_to_s1: bad arguments");
  const int64_t P = (int64_t)N * (N + 1) / 2, n = P * (P + 1) / 2;
  const unsigned blocks = (unsigned)((n + 255) / 256 < 1048576 ? (n + 255) / 256 : 1048576);
  hipStream_t st = (hipStream_t)stream;
  if (dtype_in == PSCF_F64 && dtype_out == PSCF_F64)
    s8_to_s1_kernel<double, double><<<blocks, 256, 0, st>>>(N, n, (const double*)d_s8, (double*)d_s1);
  else if (dtype_in == PSCF_F64)
    s8_to_s1_kernel<double, float><<<blocks, 256, 0, st>>>(N, n, (const double*)d_s8, (float*)d_s1);
  else if (dtype_out == PSCF_F64)
    s8_to_s1_kernel<float, double><<<blocks, 256, 0, st>>>(N, n, (const float*)d_s8, (double*)d_s1);
  else
    s8_to_s1_kernel<float, float><<<blocks, 256, 0, st>>>(N, n, (const float*)d_s8, (float*)d_s1);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_jk_s1(int dtype, const void* d_s1, int N, const double* d_dm, double hyb, double* d_out,
                          void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_jk_s1: bad dtype");
  PSCF_REQUIRE(N > 0 && N <= 4096 && d_s1 && d_dm && d_out, "pscf_jk_s1: bad arguments");
  const size_t lds = sizeof(double) * 3 * (size_t)N;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64)
    jk_s1_kernel<double><<<N, 256, lds, st>>>(N, (const double*)d_s1, d_dm, hyb / 2.0, d_out);
  else
    jk_s1_kernel<float><<<N, 256, lds, st>>>(N, (const float*)d_s1, d_dm, hyb / 2.0, d_out);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_xc_large_parts(int g) { return (g + 255) / 256; }

extern "C" int pscf_xc_large_stage(int stage, int ao_type, int N, int g, const double* d_X, const void* d_phi,
                                   const double* d_w, double eps, double* d_coef, double* d_e_part, void* d_Y,
                                   void* stream) {
  PSCF_REQUIRE(stage == 0 || stage == 1, "pscf_xc_large_stage: bad stage");
  PSCF_REQUIRE(N > 0 && g > 0 && d_phi && d_coef, "pscf_xc_large_stage: bad arguments");
  PSCF_REQUIRE(stage == 1 || (d_X && d_w && d_e_part), "pscf_xc_large_stage: stage 0 needs X, w, e_part");
  PSCF_REQUIRE(stage == 0 || d_Y, "pscf_xc_large_stage: stage 1 needs Y");
  hipStream_t st = (hipStream_t)stream;
  switch (ao_type) {
    case L_BF16: return xc_large_launch<L_BF16>(stage, N, g, d_X, d_phi, d_w, eps, d_coef, d_e_part, d_Y, st);
    case L_F16: return xc_large_launch<L_F16>(stage, N, g, d_X, d_phi, d_w, eps, d_coef, d_e_part, d_Y, st);
    case L_F32: return xc_large_launch<L_F32>(stage, N, g, d_X, d_phi, d_w, eps, d_coef, d_e_part, d_Y, st);
    case L_F64: return xc_large_launch<L_F64>(stage, N, g, d_X, d_phi, d_w, eps, d_coef, d_e_part, d_Y, st);
    default: break;
  }
  PSCF_REQUIRE(false, "pscf_xc_large_stage: bad AO type %d", ao_type);
  return PSCF_EINVAL;
}

extern "C" int pscf_xc_large_finish(int N, const double* d_Vh, const double* d_e_part, int nparts, double* d_vxc,
                                    double* d_exc, void* stream) {
  PSCF_REQUIRE(N > 0 && nparts > 0 && d_Vh && d_e_part && d_vxc && d_exc, "pscf_xc_large_finish: bad arguments");
  xc_finish_large_kernel<<<(N * N + 255) / 256, 256, 0, (hipStream_t)stream>>>(N, d_Vh, d_e_part, nparts, d_vxc,
                                                                                d_exc);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}
