// SCF step (DIIS + generalized eigensolve + density) and energy on gfx950.
//
// Reference: nanoDFT_iteration (pyscf_ipu/nanoDFT/nanoDFT.py:40-102): H = H_core
// + dJK + V_xc; DIIS (:338-364): errvec = S D H - (S D H)^T into a 9-slot ring,
// B[head+1, 1:] = <V_k, errvec>, c = pinv(masked B)[0], H = sum c_k H_k;
// eigvects = L_inv^T eigh(L_inv H L_inv^T)[1] (:69, linalg_eigh :370-393);
// D = 2 C_occ C_occ^T (:72); energy (:20-38).
// One workgroup does the whole step in FP64 with the working matrices in LDS.
// The eigensolver is a parallel cyclic Jacobi (round-robin pairing, N/2
// disjoint rotations per round), warm-started from the previous step's
// eigenvectors: A' = U_prev^T A U_prev is nearly diagonal after the first SCF
// iterations, so a few sweeps suffice.
#include "common.h"

namespace pscf {

constexpr int SCF_THREADS = 256;
constexpr int DIIS_N = 9;
constexpr int MAX_SWEEPS = 30;

// state layout (doubles): Vhist[9][N*N] | Hhist[9][N*N] | B[10][10] | Uprev[N*N] | flags[2]
__host__ __device__ inline size_t state_doubles(int N) {
  return (size_t)2 * DIIS_N * N * N + 100 + (size_t)N * N + 2;
}

// C (LDS or global, ldc = N) = op(A) * op(B), N x N, all FP64, block-cooperative.
// TA / TB: use the transpose.
template <bool TA, bool TB>
__device__ inline void matmul(const double* __restrict__ A, const double* __restrict__ B, double* __restrict__ C,
                              int N) {
  const int NN = N * N;
  for (int x = threadIdx.x; x < NN; x += blockDim.x) {
    const int i = x / N, j = x - i * N;
    double s = 0.0;
    for (int k = 0; k < N; ++k) {
      const double a = TA ? A[k * N + i] : A[i * N + k];
      const double b = TB ? B[j * N + k] : B[k * N + j];
      s += a * b;
    }
    C[x] = s;
  }
}

__device__ inline double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[k];
  __syncthreads();
  return s;
}

// Single-thread cyclic Jacobi eigen-decomposition of a small symmetric matrix.
__device__ inline void small_jacobi(double* a, double* v, int n) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) v[i * n + j] = (i == j);
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        tot += a[i * n + j] * a[i * n + j];
        if (i != j) off += a[i * n + j] * a[i * n + j];
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[p * n + q];
        if (fabs(apq) < 1e-300) continue;
        const double th = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = a[k * n + p], akq = a[k * n + q];
          a[k * n + p] = c * akp - s * akq;
          a[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = a[p * n + k], aqk = a[q * n + k];
          a[p * n + k] = c * apk - s * aqk;
          a[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = v[k * n + p], vkq = v[k * n + q];
          v[k * n + p] = c * vkp - s * vkq;
          v[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
}

// Parallel cyclic Jacobi on A (LDS, N x N, destroyed: diagonal = eigenvalues),
// accumulating rotations into V (LDS, N x N; caller initialises it).
__device__ inline void parallel_jacobi(double* A, double* V, int N, int* rp, int* rq, double* rc, double* rs,
                                       double* red) {
  const int m = N + (N & 1);  // even number of players, index N is a dummy
  const int half = m / 2;
  for (int sweep = 0; sweep < MAX_SWEEPS; ++sweep) {
    // convergence test: off-diagonal vs total Frobenius norm
    double off = 0.0, tot = 0.0;
    for (int x = threadIdx.x; x < N * N; x += blockDim.x) {
      const int i = x / N, j = x - i * N;
      const double v = A[x] * A[x];
      tot += v;
      if (i != j) off += v;
    }
    off = block_sum(off, red);
    tot = block_sum(tot, red);
    if (off <= 1e-26 * tot) break;
    for (int r = 0; r < m - 1; ++r) {
      if ((int)threadIdx.x < half) {
        const int k = threadIdx.x;
        int p, q;
        if (k == 0) {
          p = m - 1;
          q = r;
        } else {
          p = (r + k) % (m - 1);
          q = (r - k + (m - 1)) % (m - 1);
        }
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double c = 1.0, s = 0.0;
        if (q < N) {
          const double apq = A[p * N + q];
          const double app = A[p * N + p], aqq = A[q * N + q];
          if (fabs(apq) > 1e-300 && fabs(apq) > 1e-18 * sqrt(fabs(app * aqq))) {
            const double th = (aqq - app) / (2.0 * apq);
            const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            s = t * c;
          }
        }
        rp[k] = p;
        rq[k] = q;
        rc[k] = c;
        rs[k] = s;
      }
      __syncthreads();
      // rows: A <- J^T A
      for (int x = threadIdx.x; x < half * N; x += blockDim.x) {
        const int k = x / N, j = x - k * N;
        const int p = rp[k], q = rq[k];
        if (q >= N || rs[k] == 0.0) continue;
        const double c = rc[k], s = rs[k];
        const double apj = A[p * N + j], aqj = A[q * N + j];
        A[p * N + j] = c * apj - s * aqj;
        A[q * N + j] = s * apj + c * aqj;
      }
      __syncthreads();
      // columns: A <- A J, V <- V J
      for (int x = threadIdx.x; x < half * N; x += blockDim.x) {
        const int k = x / N, i = x - k * N;
        const int p = rp[k], q = rq[k];
        if (q >= N || rs[k] == 0.0) continue;
        const double c = rc[k], s = rs[k];
        const double aip = A[i * N + p], aiq = A[i * N + q];
        A[i * N + p] = c * aip - s * aiq;
        A[i * N + q] = s * aip + c * aiq;
        const double vip = V[i * N + p], viq = V[i * N + q];
        V[i * N + p] = c * vip - s * viq;
        V[i * N + q] = s * vip + c * viq;
      }
      __syncthreads();
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(SCF_THREADS) scf_step_kernel(int it, int N, int nocc, const double* __restrict__ Hcore,
                                                               const double* __restrict__ S,
                                                               const double* __restrict__ Linv,
                                                               const T* __restrict__ dJK, const T* __restrict__ Vxc,
                                                               double* __restrict__ D, double* __restrict__ Hout,
                                                               double* __restrict__ mo_e, double* __restrict__ Cout,
                                                               double* __restrict__ state, int use_diis) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int NN = N * N;
  double* A = reinterpret_cast<double*>(smem);  // N*N
  double* U = A + NN;                            // N*N
  double* Bm = U + NN;                           // 10*10 work
  double* Bv = Bm + 100;                         // 10*10 eigenvectors
  double* cvec = Bv + 100;                       // 10
  double* red = cvec + 16;                       // 8
  double* rc = red + 8;                          // N/2+1
  double* rs = rc + (N / 2 + 1);
  int* rp = reinterpret_cast<int*>(rs + (N / 2 + 1));
  int* rq = rp + (N / 2 + 1);
  int* rank = rq + (N / 2 + 1);                  // N

  double* Vh = state;
  double* Hh = state + (size_t)DIIS_N * NN;
  double* Bst = Hh + (size_t)DIIS_N * NN;
  double* Uprev = Bst + 100;
  double* flags = Uprev + NN;
  const int tid = threadIdx.x;

  // H = H_core + dJK + V_xc
  for (int x = tid; x < NN; x += blockDim.x) A[x] = Hcore[x] + (double)dJK[x] + (double)Vxc[x];
  __syncthreads();

  if (use_diis) {
    const int head = it % DIIS_N;
    if (it == 0) {
      for (int x = tid; x < 100; x += blockDim.x) {
        const int i = x / 10, j = x % 10;
        Bst[x] = ((i == 0) != (j == 0)) ? 1.0 : 0.0;
      }
    }
    // U = D H ; then errvec = S U - (S U)^T written to Vh[head]; H to Hh[head]
    matmul<false, false>(D, A, U, N);
    __syncthreads();
    double* err = Vh + (size_t)head * NN;
    for (int x = tid; x < NN; x += blockDim.x) {
      const int i = x / N, j = x - i * N;
      double sij = 0.0, sji = 0.0;
      for (int k = 0; k < N; ++k) {
        sij += S[i * N + k] * U[k * N + j];
        sji += S[j * N + k] * U[k * N + i];
      }
      err[x] = sij - sji;
      Hh[(size_t)head * NN + x] = A[x];
    }
    __threadfence_block();
    __syncthreads();
    const int nvalid = it + 1 < DIIS_N ? it + 1 : DIIS_N;
    for (int k = 0; k < DIIS_N; ++k) {
      double part = 0.0;
      if (k < nvalid)
        for (int x = tid; x < NN; x += blockDim.x) part += Vh[(size_t)k * NN + x] * err[x];
      const double dot = block_sum(part, red);
      if (tid == 0) {
        Bst[(head + 1) * 10 + (k + 1)] = dot;
        Bst[(k + 1) * 10 + (head + 1)] = dot;
      }
    }
    __syncthreads();
    if (tid == 0) {
      double mask[10];
      mask[0] = 1.0;
      for (int k = 0; k < DIIS_N; ++k) mask[k + 1] = k < nvalid ? 1.0 : 0.0;
      for (int i = 0; i < 10; ++i)
        for (int j = 0; j < 10; ++j) Bm[i * 10 + j] = Bst[i * 10 + j] * mask[i] * mask[j];
      small_jacobi(Bm, Bv, 10);
      double wmax = 0.0;
      for (int i = 0; i < 10; ++i) wmax = fmax(wmax, fabs(Bm[i * 10 + i]));
      const double cut = 10.0 * 10.0 * 2.220446049250313e-16 * wmax;
      // pinv row 0: c_j = sum_k V[0,k] V[j,k] / w_k over |w_k| > cut
      for (int j = 0; j < 10; ++j) {
        double s = 0.0;
        for (int k = 0; k < 10; ++k) {
          const double wk = Bm[k * 10 + k];
          if (fabs(wk) > cut) s += Bv[0 * 10 + k] * Bv[j * 10 + k] / wk;
        }
        cvec[j] = s;
      }
    }
    __syncthreads();
    for (int x = tid; x < NN; x += blockDim.x) {
      double h = 0.0;
      for (int k = 0; k < DIIS_N; ++k) h += cvec[k + 1] * Hh[(size_t)k * NN + x];
      A[x] = h;
    }
    __syncthreads();
  }
  for (int x = tid; x < NN; x += blockDim.x) Hout[x] = A[x];
  // M = L_inv H L_inv^T
  matmul<false, true>(A, Linv, U, N);  // U = H L^T
  __syncthreads();
  matmul<false, false>(Linv, U, A, N);  // A = L U
  __syncthreads();
  const bool warm = flags[0] != 0.0;
  if (warm) {
    // A <- Uprev^T A Uprev ; V starts at Uprev
    matmul<false, false>(A, Uprev, U, N);  // U = A Uprev
    __syncthreads();
    matmul<true, false>(Uprev, U, A, N);  // A = Uprev^T U
    __syncthreads();
    for (int x = tid; x < NN; x += blockDim.x) U[x] = Uprev[x];
  } else {
    for (int x = tid; x < NN; x += blockDim.x) U[x] = (x / N == x % N) ? 1.0 : 0.0;
  }
  __syncthreads();
  parallel_jacobi(A, U, N, rp, rq, rc, rs, red);
  // ranks of eigenvalues (ascending, ties by index)
  for (int j = tid; j < N; j += blockDim.x) {
    const double lj = A[j * N + j];
    int r = 0;
    for (int k = 0; k < N; ++k) {
      const double lk = A[k * N + k];
      r += (lk < lj) || (lk == lj && k < j);
    }
    rank[j] = r;
  }
  __syncthreads();
  for (int j = tid; j < N; j += blockDim.x) mo_e[rank[j]] = A[j * N + j];
  // save eigenvectors of the orthogonalised problem for the next warm start
  for (int x = tid; x < NN; x += blockDim.x) Uprev[x] = U[x];
  if (tid == 0) flags[0] = 1.0;
  __syncthreads();
  // C = L_inv^T U with columns permuted to ascending eigenvalue order (into A)
  for (int x = tid; x < NN; x += blockDim.x) {
    const int i = x / N, j = x - i * N;
    double s = 0.0;
    for (int k = 0; k < N; ++k) s += Linv[k * N + i] * U[k * N + j];
    A[i * N + rank[j]] = s;
  }
  __syncthreads();
  for (int x = tid; x < NN; x += blockDim.x) Cout[x] = A[x];
  // D = 2 C_occ C_occ^T
  for (int x = tid; x < NN; x += blockDim.x) {
    const int i = x / N, j = x - i * N;
    double s = 0.0;
    for (int k = 0; k < nocc; ++k) s += A[i * N + k] * A[j * N + k];
    D[x] = 2.0 * s;
  }
}

template <typename T>
__global__ void energy_kernel(int N, const double* __restrict__ D, const double* __restrict__ Hcore,
                              const T* __restrict__ dJK, const double* __restrict__ exc, double e_nuc,
                              double* __restrict__ out) {
  __shared__ double red[16];
  double s1 = 0.0, c1 = 0.0, s2 = 0.0, c2 = 0.0;  // Neumaier sums
  for (int x = threadIdx.x; x < N * N; x += blockDim.x) {
    const double a = D[x] * Hcore[x];
    double t = s1 + a;
    c1 += fabs(s1) >= fabs(a) ? (s1 - t) + a : (a - t) + s1;
    s1 = t;
    const double b = D[x] * (double)dJK[x];
    t = s2 + b;
    c2 += fabs(s2) >= fabs(b) ? (s2 - t) + b : (b - t) + s2;
    s2 = t;
  }
  const double e_core = block_sum(s1 + c1, red);
  const double e_jk = 0.5 * block_sum(s2 + c2, red);
  if (threadIdx.x == 0) {
    const double ex = *exc;
    out[0] = e_core + e_jk + ex + e_nuc;
    out[1] = e_core;
    out[2] = e_jk;
    out[3] = ex;
    out[4] = e_nuc;
  }
}

static size_t scf_lds(int N) {
  const int h = N / 2 + 1;
  return sizeof(double) * (2 * (size_t)N * N + 200 + 16 + 8 + 2 * h) + sizeof(int) * (2 * h + N);
}

}  // namespace pscf

using namespace pscf;

extern "C" size_t pscf_scf_state_bytes(int N) { return sizeof(double) * state_doubles(N); }

extern "C" int pscf_scf_step(int dtype, int iteration, int N, int nocc, const double* d_Hcore, const double* d_S,
                             const double* d_Linv, const void* d_dJK, const void* d_Vxc, double* d_D, double* d_H,
                             double* d_mo_energy, double* d_C, void* d_state, int use_diis, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_scf_step: bad dtype");
  PSCF_REQUIRE(N > 1 && nocc >= 0 && nocc <= N && iteration >= 0, "pscf_scf_step: bad sizes");
  PSCF_REQUIRE(d_Hcore && d_S && d_Linv && d_dJK && d_Vxc && d_D && d_H && d_mo_energy && d_C && d_state,
               "pscf_scf_step: null pointer");
  const size_t lds = scf_lds(N);
  int dev = 0, maxlds = 65536;
  PSCF_HIP(hipGetDevice(&dev));
  PSCF_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  PSCF_REQUIRE((int)lds <= maxlds, "pscf_scf_step: N=%d needs %zu B LDS (max %d)", N, lds, maxlds);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64) {
    if (lds > 65536)
      PSCF_HIP(hipFuncSetAttribute((const void*)scf_step_kernel<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
    scf_step_kernel<double><<<1, SCF_THREADS, lds, st>>>(iteration, N, nocc, d_Hcore, d_S, d_Linv,
                                                         (const double*)d_dJK, (const double*)d_Vxc, d_D, d_H,
                                                         d_mo_energy, d_C, (double*)d_state, use_diis);
  } else {
    if (lds > 65536)
      PSCF_HIP(hipFuncSetAttribute((const void*)scf_step_kernel<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
    scf_step_kernel<float><<<1, SCF_THREADS, lds, st>>>(iteration, N, nocc, d_Hcore, d_S, d_Linv, (const float*)d_dJK,
                                                        (const float*)d_Vxc, d_D, d_H, d_mo_energy, d_C,
                                                        (double*)d_state, use_diis);
  }
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_energy(int dtype, int N, const double* d_D, const double* d_Hcore, const void* d_dJK,
                           const double* d_exc, double e_nuc, double* d_out, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_energy: bad dtype");
  PSCF_REQUIRE(N > 0 && d_D && d_Hcore && d_dJK && d_exc && d_out, "pscf_energy: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64)
    energy_kernel<double><<<1, 256, 0, st>>>(N, d_D, d_Hcore, (const double*)d_dJK, d_exc, e_nuc, d_out);
  else
    energy_kernel<float><<<1, 256, 0, st>>>(N, d_D, d_Hcore, (const float*)d_dJK, d_exc, e_nuc, d_out);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}
