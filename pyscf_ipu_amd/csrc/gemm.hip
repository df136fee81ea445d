// General MFMA GEMM for the precision study of the AO-grid products:
// C = op(A) op(B), row-major, inputs stored as BF16 / FP16 / FP32 / FP64 and
// accumulated in FP32 (v_mfma_f32_16x16x16_{bf16,f16}, v_mfma_f32_16x16x4_f32)
// or FP64 (inputs widened exactly to FP64, v_mfma_f64_16x16x4_f64).  The
// result is stored (or atomically added, for K split across workgroups) in FP64.
//
// Workgroup: 64 x 64 tile of C, 4 wavefronts of 32 x 32 (2 x 2 MFMA blocks of
// 16 x 16), K staged through LDS in chunks of 16.  Operand layout (lane l):
// 16x16x16 types: A[l % 16][4 (l / 16) + j], B[4 (l / 16) + j][l % 16];
// 16x16x4 types: A[l % 16][l / 16], B[l / 16][l % 16]; accumulator register j
// holds row 4 (l / 16) + j (FP32) or l / 16 + 4 j (FP64), column l % 16.
#include <hip/hip_fp16.h>

#include <type_traits>

#include "common.h"

namespace pscf {

typedef short gs16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 gh16x4 __attribute__((ext_vector_type(4)));
typedef float gf32x4 __attribute__((ext_vector_type(4)));
typedef double gf64x4 __attribute__((ext_vector_type(4)));

enum GemmType : int { G_BF16 = 0, G_F16 = 1, G_F32 = 2, G_F64 = 3 };

template <int TY>
struct GIn;
template <>
struct GIn<G_BF16> {
  typedef unsigned short t;
  static __device__ inline float f(t v) { return __uint_as_float(((unsigned)v) << 16); }
};
template <>
struct GIn<G_F16> {
  typedef _Float16 t;
  static __device__ inline float f(t v) { return (float)v; }
};
template <>
struct GIn<G_F32> {
  typedef float t;
  static __device__ inline float f(t v) { return v; }
};
template <>
struct GIn<G_F64> {
  typedef double t;
  static __device__ inline double f(t v) { return v; }
};

constexpr int QM = 64, QN = 64, QK = 16;

// ACC64: accumulate in FP64; otherwise FP32 with the type's own MFMA.
template <int TY, bool ACC64, bool TA, bool TB, bool ATOMIC>
__global__ void __launch_bounds__(256) gemm_kernel(int M, int N, int K, const typename GIn<TY>::t* __restrict__ A,
                                                   int64_t lda, const typename GIn<TY>::t* __restrict__ B, int64_t ldb,
                                                   double* __restrict__ C, int64_t ldc) {
  typedef typename GIn<TY>::t Tin;
  // staged tiles, already widened to the compute type of the MFMA inputs
  typedef typename std::conditional<ACC64, double, typename std::conditional<TY == G_F64, double, float>::type>::type
      Tst;
  __shared__ Tst As[QM][QK + 1];
  __shared__ Tst Bs[QK][QN + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int m0 = blockIdx.y * QM, n0 = blockIdx.x * QN;
  const int kchunk = ((K + gridDim.z - 1) / gridDim.z + QK - 1) / QK * QK;
  const int kbeg = blockIdx.z * kchunk, kend = min(K, kbeg + kchunk);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  typedef typename std::conditional<ACC64, gf64x4, gf32x4>::type Acc;
  Acc acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[i][j][q] = 0;
  for (int k0 = kbeg; k0 < kend; k0 += QK) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = tid + r * 256;
      const int am = e / QK, ak = e % QK;
      const int gm = m0 + am, gk = k0 + ak;
      Tst va = 0;
      if (gm < M && gk < kend) va = (Tst)GIn<TY>::f(TA ? A[(int64_t)gk * lda + gm] : A[(int64_t)gm * lda + gk]);
      As[am][ak] = va;
      const int bk = e / QN, bn = e % QN;
      const int gk2 = k0 + bk, gn = n0 + bn;
      Tst vb = 0;
      if (gn < N && gk2 < kend) vb = (Tst)GIn<TY>::f(TB ? B[(int64_t)gn * ldb + gk2] : B[(int64_t)gk2 * ldb + gn]);
      Bs[bk][bn] = vb;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int rm = wm + i * 16 + li, cn = wn + j * 16 + li;
        if constexpr (ACC64) {
#pragma unroll
          for (int kk = 0; kk < QK; kk += 4)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)As[rm][kk + lg], (double)Bs[kk + lg][cn],
                                                             acc[i][j], 0, 0, 0);
        } else if constexpr (TY == G_F32 || TY == G_F64) {
#pragma unroll
          for (int kk = 0; kk < QK; kk += 4)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32((float)As[rm][kk + lg], (float)Bs[kk + lg][cn], acc[i][j],
                                                             0, 0, 0);
        } else if constexpr (TY == G_BF16) {
          gs16x4 a, b;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            a[q] = (short)(__float_as_uint((float)As[rm][4 * lg + q]) >> 16);
            b[q] = (short)(__float_as_uint((float)Bs[4 * lg + q][cn]) >> 16);
          }
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i][j], 0, 0, 0);
        } else {
          gh16x4 a, b;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            a[q] = (_Float16)(float)As[rm][4 * lg + q];
            b[q] = (_Float16)(float)Bs[4 * lg + q][cn];
          }
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, acc[i][j], 0, 0, 0);
        }
      }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int rr = ACC64 ? (lg + 4 * q) : (4 * lg + q);
        const int row = m0 + wm + i * 16 + rr, col = n0 + wn + j * 16 + li;
        if (row >= M || col >= N) continue;
        const double v = (double)acc[i][j][q];
        if (ATOMIC)
          atomicAdd(C + (int64_t)row * ldc + col, v);
        else
          C[(int64_t)row * ldc + col] = v;
      }
}

template <int TY, bool ACC64>
static int gemm_dispatch(int ta, int tb, int atomic, int M, int N, int K, const void* A, int64_t lda, const void* B,
                         int64_t ldb, double* C, int64_t ldc, int ksplit, hipStream_t st) {
  typedef typename GIn<TY>::t T;
  dim3 grid((N + QN - 1) / QN, (M + QM - 1) / QM, ksplit);
  const T* a = (const T*)A;
  const T* b = (const T*)B;
#define PSCF_GEMM_LAUNCH(TA_, TB_, AT_) \
  gemm_kernel<TY, ACC64, TA_, TB_, AT_><<<grid, 256, 0, st>>>(M, N, K, a, lda, b, ldb, C, ldc)
  if (!ta && !tb && !atomic) PSCF_GEMM_LAUNCH(false, false, false);
  else if (!ta && tb && !atomic) PSCF_GEMM_LAUNCH(false, true, false);
  else if (!ta && tb && atomic) PSCF_GEMM_LAUNCH(false, true, true);
  else if (!ta && !tb && atomic) PSCF_GEMM_LAUNCH(false, false, true);
  else if (ta && !tb && !atomic) PSCF_GEMM_LAUNCH(true, false, false);
  else if (ta && !tb && atomic) PSCF_GEMM_LAUNCH(true, false, true);
  else if (ta && tb && !atomic) PSCF_GEMM_LAUNCH(true, true, false);
  else PSCF_GEMM_LAUNCH(true, true, true);
#undef PSCF_GEMM_LAUNCH
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

}  // namespace pscf

using namespace pscf;

extern "C" int pscf_gemm(int in_type, int acc64, int trans_a, int trans_b, int atomic, int M, int N, int K,
                         const void* d_A, int64_t lda, const void* d_B, int64_t ldb, double* d_C, int64_t ldc,
                         int ksplit, void* stream) {
  PSCF_REQUIRE(in_type >= G_BF16 && in_type <= G_F64, "pscf_gemm: bad input type %d", in_type);
  PSCF_REQUIRE(M >= 0 && N >= 0 && K >= 0 && ksplit >= 1 && ksplit <= 65535, "pscf_gemm: bad sizes");
  PSCF_REQUIRE(ksplit == 1 || atomic, "pscf_gemm: ksplit > 1 needs atomic accumulation");
  PSCF_REQUIRE(M == 0 || N == 0 || (d_A && d_B && d_C), "pscf_gemm: null pointer");
  if (M == 0 || N == 0) return PSCF_OK;
  hipStream_t st = (hipStream_t)stream;
  const int ta = trans_a != 0, tb = trans_b != 0, at = atomic != 0;
  if (acc64) {
    switch (in_type) {
      case G_BF16: return gemm_dispatch<G_BF16, true>(ta, tb, at, M, N, K, d_A, lda, d_B, ldb, d_C, ldc, ksplit, st);
      case G_F16: return gemm_dispatch<G_F16, true>(ta, tb, at, M, N, K, d_A, lda, d_B, ldb, d_C, ldc, ksplit, st);
      case G_F32: return gemm_dispatch<G_F32, true>(ta, tb, at, M, N, K, d_A, lda, d_B, ldb, d_C, ldc, ksplit, st);
      default: return gemm_dispatch<G_F64, true>(ta, tb, at, M, N, K, d_A, lda, d_B, ldb, d_C, ldc, ksplit, st);
    }
  }
  switch (in_type) {
    case G_BF16: return gemm_dispatch<G_BF16, false>(ta, tb, at, M, N, K, d_A, lda, d_B, ldb, d_C, ldc, ksplit, st);
    case G_F16: return gemm_dispatch<G_F16, false>(ta, tb, at, M, N, K, d_A, lda, d_B, ldb, d_C, ldc, ksplit, st);
    case G_F32: return gemm_dispatch<G_F32, false>(ta, tb, at, M, N, K, d_A, lda, d_B, ldb, d_C, ldc, ksplit, st);
    default: return gemm_dispatch<G_F64, false>(ta, tb, at, M, N, K, d_A, lda, d_B, ldb, d_C, ldc, ksplit, st);
  }
}
