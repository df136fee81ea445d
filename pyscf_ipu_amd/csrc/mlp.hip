// Learned initial guess on gfx950: atom-pair features, an MLP trained with
// BF16 MFMA (FP32 accumulation, FP32 master weights, Adam), and assembly of
// the predicted STO-3G density blocks into per-molecule density matrices.
//
// There is no counterpart in the reference (its SCF always starts from PySCF's
// MINAO guess, pyscf_ipu/nanoDFT/nanoDFT.py:233).  Network: per ordered atom
// pair, 43 inputs (one-hot H/C/N/O of both atoms, 32 Gaussian RBFs of the
// distance on 0..5 Angstrom with sigma 0.15 Angstrom, unit direction) ->
// 256 -> 256 -> 256 (SiLU) -> 25 (the 5x5 block of AOs of atom a x AOs of b).
#include <string.h>

#include "batch.h"
#include "common.h"

namespace pscf {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;

__host__ __device__ inline bf16_t f2bf(float x) {  // round to nearest even
#ifdef __HIP_DEVICE_COMPILE__
  unsigned u = __float_as_uint(x);
#else
  unsigned u;
  memcpy(&u, &x, 4);
#endif
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ inline float bf2f(bf16_t h) { return __uint_as_float(((unsigned)h) << 16); }
__device__ inline float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ inline float dsilu(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

enum Epilogue : int {
  EPI_BIAS_SILU = 0,  // Z = acc + b (FP32 out), H = silu(Z) (BF16 out2)
  EPI_BIAS = 1,       // out = acc + b (FP32)
  EPI_DSILU = 2,      // dZ = acc * dsilu(Zprev) -> FP32 out and BF16 out2
  EPI_ATOMIC = 3,     // out += acc (FP32 atomics; split-K)
};

constexpr int GBM = 64, GBN = 64, GBK = 16;

// C[M][N] = op(A)[M][K] * op(B)[K][N], BF16 operands, FP32 accumulation on
// v_mfma_f32_16x16x16bf16_1k.  A(m,k) = TA ? A[k*lda+m] : A[m*lda+k];
// B(k,n) = TB ? B[n*ldb+k] : B[k*ldb+n].  Workgroup: 64x64 tile, 4 waves of
// 32x32 (2x2 MFMA blocks).  blockIdx.z splits K into gridDim.z ranges (only
// with EPI_ATOMIC).  Out-of-range rows/cols/ks are zero-filled.
template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(int M, int N, int K, const bf16_t* __restrict__ A, int lda,
                                                        const bf16_t* __restrict__ B, int ldb,
                                                        const float* __restrict__ bias, const float* __restrict__ aux,
                                                        float* __restrict__ out, bf16_t* __restrict__ out2, int ldc) {
  __shared__ bf16_t As[GBM][GBK + 4];
  __shared__ bf16_t Bs[GBK][GBN + 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * GBM, n0 = blockIdx.x * GBN;
  const int kchunk = (K + gridDim.z - 1) / gridDim.z;
  const int kbeg = blockIdx.z * kchunk, kend = min(K, kbeg + kchunk);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (int k0 = kbeg; k0 < kend; k0 += GBK) {
    // stage 64x16 of A and 16x64 of B (4 elements per thread each)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = tid + r * 256;
      const int am = e / GBK, ak = e % GBK;
      const int gm = m0 + am, gk = k0 + ak;
      bf16_t va = 0;
      if (gm < M && gk < kend) va = TA ? A[(int64_t)gk * lda + gm] : A[(int64_t)gm * lda + gk];
      As[am][ak] = va;
      const int bk = e / GBN, bn = e % GBN;
      const int gk2 = k0 + bk, gn = n0 + bn;
      bf16_t vb = 0;
      if (gn < N && gk2 < kend) vb = TB ? B[(int64_t)gn * ldb + gk2] : B[(int64_t)gk2 * ldb + gn];
      Bs[bk][bn] = vb;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        bf16x4 a, b;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          a[q] = (short)As[wm + i * 16 + (lane & 15)][4 * (lane >> 4) + q];
          b[q] = (short)Bs[4 * (lane >> 4) + q][wn + j * 16 + (lane & 15)];
        }
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i][j], 0, 0, 0);
      }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + wm + i * 16 + 4 * (lane >> 4) + q;
        const int col = n0 + wn + j * 16 + (lane & 15);
        if (row >= M || col >= N) continue;
        const float v = acc[i][j][q];
        const int64_t o = (int64_t)row * ldc + col;
        if (EPI == EPI_BIAS_SILU) {
          const float z = v + bias[col];
          out[o] = z;
          out2[o] = f2bf(silu(z));
        } else if (EPI == EPI_BIAS) {
          out[o] = v + bias[col];
        } else if (EPI == EPI_DSILU) {
          const float dz = v * dsilu(aux[o]);
          out[o] = dz;
          out2[o] = f2bf(dz);
        } else {
          atomicAdd(out + o, v);
        }
      }
}

// g = 2 (pred - Y) mask / nvalid on the padded output (ld = 32); BF16 and
// FP32 copies of g; loss accumulated into *loss (FP64).
__global__ void mse_grad_kernel(int B, const float* __restrict__ pred, const float* __restrict__ Y,
                                const float* __restrict__ mask, float inv_n, float* __restrict__ g32,
                                bf16_t* __restrict__ g16, double* __restrict__ loss) {
  __shared__ double red[4];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (i < (int64_t)B * 32) {
    const float d = (pred[i] - Y[i]) * mask[i];
    l = (double)d * d * inv_n;
    const float g = 2.0f * d * inv_n;
    g32[i] = g;
    g16[i] = f2bf(g);
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, red[0] + red[1] + red[2] + red[3]);
}

// db[n] = sum_m dZ[m][n]  (one workgroup per column, fixed-order tree)
__global__ void colsum_kernel(int M, int N, const float* __restrict__ dz, int ld, float* __restrict__ db) {
  __shared__ float red[4];
  const int n = blockIdx.x;
  float s = 0.f;
  for (int m = threadIdx.x; m < M; m += blockDim.x) s += dz[(int64_t)m * ld + n];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) db[n] = red[0] + red[1] + red[2] + red[3];
}

// Adam on FP32 master weights; BF16 working copy refreshed.
__global__ void adam_kernel(int64_t n, float* __restrict__ w, bf16_t* __restrict__ w16, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  w[i] = wi;
  if (w16) w16[i] = f2bf(wi);
}

// Pair features.  pairs[p] = (global atom a, global atom b); rows padded to 64
// BF16 columns (43 used) and an FP32 copy (ld 64) for checks.
__global__ void pair_features_kernel(int64_t P, const int* __restrict__ pairs, const int* __restrict__ atm,
                                     const double* __restrict__ env, bf16_t* __restrict__ X16,
                                     float* __restrict__ X32, unsigned char* __restrict__ valid) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int a = pairs[2 * p], b = pairs[2 * p + 1];
  const double* Ra = env + atm[a * ATM_SLOTS + 1];
  const double* Rb = env + atm[b * ATM_SLOTS + 1];
  const double bohr = 0.52917721092;
  const double vx = (Rb[0] - Ra[0]) * bohr, vy = (Rb[1] - Ra[1]) * bohr, vz = (Rb[2] - Ra[2]) * bohr;
  const double d = sqrt(vx * vx + vy * vy + vz * vz);
  float f[64];
#pragma unroll
  for (int k = 0; k < 64; ++k) f[k] = 0.f;
  const int za = atm[a * ATM_SLOTS], zb = atm[b * ATM_SLOTS];
  const int ea = za == 1 ? 0 : (za == 6 ? 1 : (za == 7 ? 2 : 3));
  const int eb = zb == 1 ? 0 : (zb == 6 ? 1 : (zb == 7 ? 2 : 3));
  f[ea] = 1.f;
  f[4 + eb] = 1.f;
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    const double mu = 5.0 * k / 31.0;
    const double t = (d - mu) / 0.15;
    f[8 + k] = (float)exp(-0.5 * t * t);
  }
  if (d > 0.0) {
    f[40] = (float)(vx / d);
    f[41] = (float)(vy / d);
    f[42] = (float)(vz / d);
  }
#pragma unroll
  for (int k = 0; k < 64; ++k) {
    X16[p * 64 + k] = f2bf(f[k]);
    if (X32) X32[p * 64 + k] = f[k];
  }
  valid[p] = d < 5.0 ? 1 : 0;
}

// Targets: Y[p][r*5+c] = D[AO(a)+r][AO(b)+c] of the pair's molecule (ld 32),
// mask 1 where r < nao(a), c < nao(b) and the pair is within the cutoff.
__global__ void pair_targets_kernel(int64_t P, const int* __restrict__ pinfo, const unsigned char* __restrict__ valid,
                                    const BatchMol* __restrict__ mols, const double* __restrict__ D,
                                    float* __restrict__ Y, float* __restrict__ mask) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * 32) return;
  const int64_t p = t >> 5;
  const int e = (int)(t & 31);
  const int* q = pinfo + 5 * p;  // mol, ao_a, n_a, ao_b, n_b
  const int r = e / 5, c = e % 5;
  float y = 0.f, mk = 0.f;
  if (e < 25 && r < q[2] && c < q[4] && valid[p]) {
    const BatchMol Mo = mols[q[0]];
    y = (float)D[Mo.mat0 + (int64_t)(q[1] + r) * Mo.nao + q[3] + c];
    mk = 1.f;
  }
  Y[t] = y;
  mask[t] = mk;
}

// Predicted blocks into D (zero-initialised by the caller); pairs beyond the
// cutoff contribute nothing.
__global__ void pair_assemble_kernel(int64_t P, const int* __restrict__ pinfo, const unsigned char* __restrict__ valid,
                                     const BatchMol* __restrict__ mols, const float* __restrict__ pred,
                                     double* __restrict__ D) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * 32) return;
  const int64_t p = t >> 5;
  const int e = (int)(t & 31);
  const int* q = pinfo + 5 * p;
  const int r = e / 5, c = e % 5;
  if (e < 25 && r < q[2] && c < q[4] && valid[p]) {
    const BatchMol Mo = mols[q[0]];
    D[Mo.mat0 + (int64_t)(q[1] + r) * Mo.nao + q[3] + c] = (double)pred[t];
  }
}

__global__ void f32_to_bf16_kernel(int64_t n, const float* __restrict__ x, bf16_t* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

template <bool TA, bool TB, int EPI>
static int launch_gemm(int M, int N, int K, const void* A, int lda, const void* B, int ldb, const float* bias,
                       const float* aux, float* out, void* out2, int ldc, int ksplit, hipStream_t st) {
  dim3 grid((N + GBN - 1) / GBN, (M + GBM - 1) / GBM, ksplit);
  gemm_bf16_kernel<TA, TB, EPI><<<grid, 256, 0, st>>>(M, N, K, (const bf16_t*)A, lda, (const bf16_t*)B, ldb, bias,
                                                      aux, out, (bf16_t*)out2, ldc);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

}  // namespace pscf

using namespace pscf;

// op: 0 = forward hidden (X W + b, SiLU; out = Z FP32, out2 = H BF16)
//     1 = forward output (X W + b; out FP32)
//     2 = backward data (dY W^T, times dsilu(aux = Z of the layer below); out FP32, out2 BF16)
//     3 = weight gradient (X^T dY accumulated into out with atomics; ksplit > 1 allowed)
extern "C" int pscf_mlp_gemm(int op, int M, int N, int K, const void* d_A, int lda, const void* d_B, int ldb,
                             const float* d_bias, const float* d_aux, float* d_out, void* d_out2, int ldc, int ksplit,
                             void* stream) {
  PSCF_REQUIRE(M >= 0 && N >= 0 && K >= 0 && d_A && d_B && d_out, "pscf_mlp_gemm: bad arguments");
  PSCF_REQUIRE(ksplit >= 1 && (ksplit == 1 || op == 3), "pscf_mlp_gemm: ksplit > 1 only for weight gradients");
  if (M == 0 || N == 0) return PSCF_OK;
  hipStream_t st = (hipStream_t)stream;
  switch (op) {
    case 0:
      PSCF_REQUIRE(d_bias && d_out2, "pscf_mlp_gemm: forward needs bias and BF16 output");
      return launch_gemm<false, false, EPI_BIAS_SILU>(M, N, K, d_A, lda, d_B, ldb, d_bias, nullptr, d_out, d_out2, ldc,
                                                      1, st);
    case 1:
      PSCF_REQUIRE(d_bias, "pscf_mlp_gemm: forward needs bias");
      return launch_gemm<false, false, EPI_BIAS>(M, N, K, d_A, lda, d_B, ldb, d_bias, nullptr, d_out, nullptr, ldc, 1,
                                                 st);
    case 2:
      PSCF_REQUIRE(d_aux && d_out2, "pscf_mlp_gemm: backward needs Z and BF16 output");
      return launch_gemm<false, true, EPI_DSILU>(M, N, K, d_A, lda, d_B, ldb, nullptr, d_aux, d_out, d_out2, ldc, 1,
                                                 st);
    case 3:
      return launch_gemm<true, false, EPI_ATOMIC>(M, N, K, d_A, lda, d_B, ldb, nullptr, nullptr, d_out, nullptr, ldc,
                                                  ksplit, st);
    default:
      break;
  }
  PSCF_REQUIRE(false, "pscf_mlp_gemm: bad op %d", op);
  return PSCF_EINVAL;
}

extern "C" int pscf_mlp_mse_grad(int B, const float* d_pred, const float* d_Y, const float* d_mask, float inv_n,
                                 float* d_g32, void* d_g16, double* d_loss, void* stream) {
  PSCF_REQUIRE(B >= 0 && d_pred && d_Y && d_mask && d_g32 && d_g16 && d_loss, "pscf_mlp_mse_grad: bad arguments");
  if (B == 0) return PSCF_OK;
  const int64_t n = (int64_t)B * 32;
  mse_grad_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(B, d_pred, d_Y, d_mask, inv_n, d_g32,
                                                                              (bf16_t*)d_g16, d_loss);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_mlp_colsum(int M, int N, const float* d_dz, int ld, float* d_db, void* stream) {
  PSCF_REQUIRE(M >= 0 && N > 0 && d_dz && d_db, "pscf_mlp_colsum: bad arguments");
  colsum_kernel<<<N, 256, 0, (hipStream_t)stream>>>(M, N, d_dz, ld, d_db);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_mlp_adam(int64_t n, float* d_w, void* d_w16, const float* d_g, float* d_m, float* d_v, float lr,
                             float b1, float b2, float eps, int step, void* stream) {
  PSCF_REQUIRE(n >= 0 && step >= 1 && d_w && d_g && d_m && d_v, "pscf_mlp_adam: bad arguments");
  if (n == 0) return PSCF_OK;
  const float c1 = 1.f / (1.f - powf(b1, (float)step)), c2 = 1.f / (1.f - powf(b2, (float)step));
  adam_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(n, d_w, (bf16_t*)d_w16, d_g, d_m, d_v, lr,
                                                                          b1, b2, eps, c1, c2);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_mlp_to_bf16(int64_t n, const float* d_x, void* d_y, void* stream) {
  PSCF_REQUIRE(n >= 0 && (n == 0 || (d_x && d_y)), "pscf_mlp_to_bf16: bad arguments");
  if (n == 0) return PSCF_OK;
  f32_to_bf16_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(n, d_x, (bf16_t*)d_y);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_pair_features(int64_t P, const int* d_pairs, const int* d_atm, const double* d_env, void* d_X16,
                                  float* d_X32, unsigned char* d_valid, void* stream) {
  PSCF_REQUIRE(P >= 0 && (P == 0 || (d_pairs && d_atm && d_env && d_X16 && d_valid)), "pscf_pair_features: bad arguments");
  if (P == 0) return PSCF_OK;
  pair_features_kernel<<<(unsigned)((P + 255) / 256), 256, 0, (hipStream_t)stream>>>(P, d_pairs, d_atm, d_env,
                                                                                   (bf16_t*)d_X16, d_X32, d_valid);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_pair_targets(int64_t P, const int* d_pinfo, const unsigned char* d_valid, const void* d_mols,
                                 const double* d_D, float* d_Y, float* d_mask, void* stream) {
  PSCF_REQUIRE(P >= 0 && (P == 0 || (d_pinfo && d_valid && d_mols && d_D && d_Y && d_mask)),
               "pscf_pair_targets: bad arguments");
  if (P == 0) return PSCF_OK;
  pair_targets_kernel<<<(unsigned)((P * 32 + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      P, d_pinfo, d_valid, (const BatchMol*)d_mols, d_D, d_Y, d_mask);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_pair_assemble(int64_t P, const int* d_pinfo, const unsigned char* d_valid, const void* d_mols,
                                  const float* d_pred, double* d_D, void* stream) {
  PSCF_REQUIRE(P >= 0 && (P == 0 || (d_pinfo && d_valid && d_mols && d_pred && d_D)),
               "pscf_pair_assemble: bad arguments");
  if (P == 0) return PSCF_OK;
  pair_assemble_kernel<<<(unsigned)((P * 32 + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      P, d_pinfo, d_valid, (const BatchMol*)d_mols, d_pred, d_D);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}
