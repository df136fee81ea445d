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
#include "batch.h"
#include "common.h"

namespace pscf {

constexpr int SCF_THREADS = 1024;
constexpr int DIIS_N = 9;
constexpr int MAX_SWEEPS = 30;

// state layout (doubles): Vhist[9][N*N] | Hhist[9][N*N] | B[10][10] | Uprev[N*N] | flags[2]
__host__ __device__ inline size_t state_doubles(int N) {
  return (size_t)2 * DIIS_N * N * N + 100 + (size_t)N * N + 2;
}

// Working matrices live in LDS with an odd leading dimension ld = N | 1, so a
// walk down a column (stride ld doubles) touches distinct LDS bank pairs.
__host__ __device__ inline int lds_ld(int N) { return N | 1; }

// C = op(X) op(Y) for N <= 64 by 16 waves on FP64 MFMA (16x16x4): wave w
// computes output tile (w / nt, w % nt), nt = ceil(N / 16).  Lane l supplies
// A[row = l % 16][k = l / 16] and B[k = l / 16][col = l % 16] and receives
// C[l / 16 + 4 j][l % 16], j = 0..3.  Operands outside N x N are zero (their
// LDS reads are clamped in bounds).  Each operand element is read once per
// tile, an eighth of the row-walk form's LDS traffic.
template <bool TX, bool TY>
__device__ inline void lds_matmul_mfma(const double* __restrict__ X, const double* __restrict__ Y,
                                       double* __restrict__ C, int N, int ld, int ldc) {
  typedef double v4d __attribute__((ext_vector_type(4)));
  const int nt = (N + 15) >> 4;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (w >= nt * nt) return;  // wave-uniform
  const int I = w / nt, J = w - I * nt;
  const int li = lane & 15, lg = lane >> 4;
  const int row = I * 16 + li, col = J * 16 + li;
  const bool rv = row < N, cv = col < N;
  const int rowc = rv ? row : N - 1, colc = cv ? col : N - 1;
  v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int k0 = 0; k0 < N; k0 += 4) {
    const int k = k0 + lg;
    const bool kv = k < N;
    const int kc = kv ? k : N - 1;
    const double a = TX ? X[kc * ld + rowc] : X[rowc * ld + kc];
    const double b = TY ? Y[colc * ld + kc] : Y[kc * ld + colc];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64((rv && kv) ? a : 0.0, (cv && kv) ? b : 0.0, acc, 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int r = I * 16 + lg + 4 * j;
    if (r < N && cv) C[r * ldc + col] = acc[j];
  }
}

// C = op(X) op(Y), N x N, X and Y in LDS (leading dimension ld), C in LDS or
// global (leading dimension ldc), never aliasing X or Y.  N <= 64 with 1024
// threads takes the MFMA form above.  Otherwise each thread owns rows
// {i, i + h} x columns {j, j + h} (h = ceil(N / 2)), so the lanes of a wave walk
// consecutive columns and share rows.
template <bool TX, bool TY>
__device__ inline void lds_matmul(const double* __restrict__ X, const double* __restrict__ Y, double* __restrict__ C,
                                  int N, int ld, int ldc) {
  if (N <= 64 && blockDim.x == 1024) {
    lds_matmul_mfma<TX, TY>(X, Y, C, N, ld, ldc);
    return;
  }
  const int h = (N + 1) >> 1;
  for (int b = threadIdx.x; b < h * h; b += blockDim.x) {
    const int i = b / h, j = b - i * h;
    const int i1 = i + h, j1 = j + h;
    const bool vi = i1 < N, vj = j1 < N;
    const int ir = vi ? i1 : i, jr = vj ? j1 : j;
    double c00 = 0.0, c01 = 0.0, c10 = 0.0, c11 = 0.0;
    for (int k = 0; k < N; ++k) {
      const double x0 = TX ? X[k * ld + i] : X[i * ld + k];
      const double x1 = TX ? X[k * ld + ir] : X[ir * ld + k];
      const double y0 = TY ? Y[j * ld + k] : Y[k * ld + j];
      const double y1 = TY ? Y[jr * ld + k] : Y[k * ld + jr];
      c00 = fma(x0, y0, c00);
      c01 = fma(x0, y1, c01);
      c10 = fma(x1, y0, c10);
      c11 = fma(x1, y1, c11);
    }
    C[i * ldc + j] = c00;
    if (vj) C[i * ldc + j1] = c01;
    if (vi) {
      C[i1 * ldc + j] = c10;
      if (vj) C[i1 * ldc + j1] = c11;
    }
  }
}

// N x N global (leading dimension N) <-> LDS (leading dimension ld)
__device__ inline void lds_load(double* __restrict__ dst, const double* __restrict__ src, int N, int ld) {
  for (int x = threadIdx.x; x < N * N; x += blockDim.x) {
    const int i = x / N;
    dst[i * ld + x - i * N] = src[x];
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

// Orders the LDS accesses of one wavefront (its lanes run in lockstep).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// 1 / sqrt(x) for x >= 1 in FP64: the hardware estimate refined by two
// Newton steps y <- y (3 - x y^2) / 2 (each doubles the correct bits).
__device__ inline double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = y * (1.5 - 0.5 * x * y * y);
  return y * (1.5 - 0.5 * x * y * y);
}

// Cyclic Jacobi eigen-decomposition of a small symmetric n x n matrix
// (n <= 16, LDS, row-major) by ONE wavefront: round-robin pairing gives n/2
// disjoint rotations per round, applied as 2x2 block updates.  On exit the
// diagonal of a holds the eigenvalues and v (LDS) the eigenvectors (columns).
__device__ inline void wave_jacobi(double* a, double* v, int n, int* rp, int* rq, double* rc, double* rs) {
  const int lane = threadIdx.x & 63;
  const int m = n + (n & 1), half = m / 2;
  for (int x = lane; x < n * n; x += 64) v[x] = (x / n == x % n) ? 1.0 : 0.0;
  wave_sync();
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int x = lane; x < n * n; x += 64) {
      const double q = a[x] * a[x];
      tot += q;
      if (x / n != x % n) off += q;
    }
    off = wave_sum(off);
    tot = wave_sum(tot);
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int r = 0; r < m - 1; ++r) {
      if (lane < half) {
        int p, q;
        if (lane == 0) {
          p = m - 1;
          q = r;
        } else {  // (r + lane) mod (m - 1), (r - lane) mod (m - 1) without a division
          p = r + lane;
          if (p >= m - 1) p -= m - 1;
          q = r - lane;
          if (q < 0) q += m - 1;
        }
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double c = 1.0, s = 0.0;
        if (q < n) {
          const double apq = a[p * n + q];
          if (fabs(apq) >= 1e-300) {  // the rotation of parallel_jacobi (one division)
            const double d = a[q * n + q] - a[p * n + p], e = 2.0 * apq;
            const double t = (d >= 0 ? e : -e) / (fabs(d) + sqrt(d * d + e * e));
            c = rsqrt_nr(1.0 + t * t);
            s = t * c;
          }
        }
        rp[lane] = p;
        rq[lane] = q;
        rc[lane] = c;
        rs[lane] = s;
      }
      wave_sync();
      for (int b = lane; b < half * half; b += 64) {
        const int k = b / half, l = b - k * half;
        const int pk = rp[k], qk = rq[k], pl = rp[l], ql = rq[l];
        const double ck = rc[k], sk = rs[k], cl = rc[l], sl = rs[l];
        const bool qkv = qk < n, qlv = ql < n;
        const double x00 = a[pk * n + pl];
        const double x01 = qlv ? a[pk * n + ql] : 0.0;
        const double x10 = qkv ? a[qk * n + pl] : 0.0;
        const double x11 = (qkv && qlv) ? a[qk * n + ql] : 0.0;
        const double t00 = cl * x00 - sl * x01, t01 = sl * x00 + cl * x01;
        const double t10 = cl * x10 - sl * x11, t11 = sl * x10 + cl * x11;
        a[pk * n + pl] = ck * t00 - sk * t10;
        if (qlv) a[pk * n + ql] = ck * t01 - sk * t11;
        if (qkv) a[qk * n + pl] = sk * t00 + ck * t10;
        if (qkv && qlv) a[qk * n + ql] = sk * t01 + ck * t11;
      }
      for (int x = lane; x < half * n; x += 64) {
        const int l = x / n, i = x - l * n;
        const int p = rp[l], q = rq[l];
        if (q >= n) continue;
        const double c = rc[l], s = rs[l];
        const double vip = v[i * n + p], viq = v[i * n + q];
        v[i * n + p] = c * vip - s * viq;
        v[i * n + q] = s * vip + c * viq;
      }
      wave_sync();
    }
  }
}

// Pair k of round r of the round-robin over m (even) players, p < q: pair 0 is
// (r, m - 1), pair k > 0 is {(r + k) mod (m - 1), (r - k) mod (m - 1)}.
__device__ inline void rr_pair(int r, int k, int m, int& p, int& q) {
  const int M1 = m - 1;
  if (k == 0) {
    p = r;
    q = M1;
    return;
  }
  int a = r + k, b = r - k;
  if (a >= M1) a -= M1;
  if (b < 0) b += M1;
  p = a < b ? a : b;
  q = a < b ? b : a;
}

// Rotation (c, s) of pair k of round r from the current A (index N is the
// dummy of odd N: identity).  t = sign(th) / (|th| + sqrt(th^2 + 1)) with
// th = d / e, d = aqq - app, e = 2 apq, written with one division:
// t = sign(d) e / (|d| + sqrt(d^2 + e^2)); entries below 1e-18 of the
// diagonal scale are left alone.
__device__ inline void jacobi_pair_rotation(const double* A, int ld, int N, int m, int r, int k, double* rc,
                                            double* rs) {
  int p, q;
  rr_pair(r, k, m, p, q);
  double c = 1.0, s = 0.0;
  if (q < N) {
    const double apq = A[p * ld + q];
    const double app = A[p * ld + p], aqq = A[q * ld + q];
    if (fabs(apq) > 1e-300 && apq * apq > 1e-36 * fabs(app * aqq)) {
      const double d = aqq - app, e = 2.0 * apq;
      const double t = (d >= 0 ? e : -e) / (fabs(d) + sqrt(d * d + e * e));
      c = rsqrt_nr(1.0 + t * t);
      s = t * c;
    }
  }
  rc[k] = c;
  rs[k] = s;
}

// One V column rotation of round r: row i of columns (p, q) of pair lp.
struct VItem {
  int a, b;  // LDS indices of V[i][p], V[i][q]
  bool on;
  double vp, vq, c, s;
};

__device__ inline VItem v_item_load(const double* V, const double* rc, const double* rs, int N, int ld, int m,
                                    int half, int r, int lp, int i) {
  VItem it;
  int p, q;
  rr_pair(r, lp < half ? lp : 0, m, p, q);
  it.on = lp < half && i < N && q < N;
  it.a = it.on ? i * ld + p : 0;
  it.b = it.on ? i * ld + q : 0;
  it.vp = V[it.a];
  it.vq = V[it.b];
  it.c = rc[lp < half ? lp : 0];
  it.s = rs[lp < half ? lp : 0];
  return it;
}

__device__ inline void v_item_store(double* V, const VItem& it) {
  if (it.on) {
    V[it.a] = it.c * it.vp - it.s * it.vq;
    V[it.b] = it.s * it.vp + it.c * it.vq;
  }
}

// Parallel cyclic Jacobi for N <= 64 by exactly 1024 threads, two barriers a
// round with the rotations off the critical path.  Round r:
//   phase 1: thread (k, l) = (tid / 32, tid % 32) updates the 2x2 block of A
//            (pair k rows, pair l columns); every wave also rotates the V
//            columns of pair w = tid / 64 (lane = row), wave 15 also pair 31;
//   phase 2: wave 15 computes round r + 1's rotations from the updated A while
//            waves 0..14 rotate the V columns of pairs 16..30 (V never feeds A).
// Rotation tables rc, rs hold two rounds (32 pairs each, by round parity).
// Pair indices come from the round-robin formula.  Every thread issues its
// loads before its stores (A blocks and V items are disjoint), so the
// arithmetic and the result are those of the two-phase form below.
__device__ inline int jacobi_pipelined(double* A, double* V, int N, int ld, double* rc, double* rs, double* red) {
  const int m = N + (N & 1), half = m / 2;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int k = tid >> 5, l = tid & 31;
  int sweep = 0;
  for (; sweep < MAX_SWEEPS; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int x = tid; x < N * N; x += blockDim.x) {
      const int i = x / N, j = x - i * N;
      const double v = A[i * ld + j] * A[i * ld + j];
      tot += v;
      if (i != j) off += v;
    }
    off = block_sum(off, red);
    tot = block_sum(tot, red);
    if (off <= 1e-18 * tot) break;
    if (tid < half) jacobi_pair_rotation(A, ld, N, m, 0, tid, rc, rs);
    __syncthreads();
    for (int r = 0; r < m - 1; ++r) {
      const int cur = (r & 1) * 32, nxt = 32 - cur;
      {  // phase 1
        const bool blk = k < half && l < half;
        int pk, qk, pl, ql;
        rr_pair(r, blk ? k : 0, m, pk, qk);
        rr_pair(r, blk ? l : 0, m, pl, ql);
        const bool qkv = blk && qk < N, qlv = blk && ql < N;
        const int rq_ = qkv ? qk : pk, cq_ = qlv ? ql : pl;
        const int i00 = pk * ld + pl, i01 = pk * ld + cq_, i10 = rq_ * ld + pl, i11 = rq_ * ld + cq_;
        const double x00 = A[i00], x01r = A[i01], x10r = A[i10], x11r = A[i11];
        const double ck = rc[cur + k], sk = rs[cur + k], cl = rc[cur + l], sl = rs[cur + l];
        const VItem v1 = v_item_load(V, rc + cur, rs + cur, N, ld, m, half, r, w, lane);
        const VItem v2 = v_item_load(V, rc + cur, rs + cur, N, ld, m, half, r, w == 15 ? 31 : 32, lane);
        const double x01 = qlv ? x01r : 0.0, x10 = qkv ? x10r : 0.0, x11 = (qkv && qlv) ? x11r : 0.0;
        // columns: [x.0, x.1] R_l; rows: R_k^T [t0.; t1.]
        const double t00 = cl * x00 - sl * x01, t01 = sl * x00 + cl * x01;
        const double t10 = cl * x10 - sl * x11, t11 = sl * x10 + cl * x11;
        if (blk) {
          A[i00] = ck * t00 - sk * t10;
          if (qlv) A[i01] = ck * t01 - sk * t11;
          if (qkv) A[i10] = sk * t00 + ck * t10;
          if (qkv && qlv) A[i11] = sk * t01 + ck * t11;
        }
        v_item_store(V, v1);
        v_item_store(V, v2);
      }
      __syncthreads();
      if (w == 15) {  // phase 2
        if (r + 1 < m - 1 && lane < half) jacobi_pair_rotation(A, ld, N, m, r + 1, lane, rc + nxt, rs + nxt);
      } else {
        const VItem v3 = v_item_load(V, rc + cur, rs + cur, N, ld, m, half, r, 16 + w, lane);
        v_item_store(V, v3);
      }
      __syncthreads();
    }
  }
  return sweep;
}

// Parallel cyclic Jacobi on A (LDS, N x N, leading dimension ld; on exit its
// diagonal holds the eigenvalues), accumulating rotations into V (LDS, same
// leading dimension, caller initialised).  Round-robin pairing gives N/2
// disjoint rotations per round; the two-sided update A <- J^T A J is one pass
// over the 2x2 blocks (pair k rows, pair l columns): Y = R_k^T X R_l with
// R = [[c, s], [-s, c]].  N <= 64 with 1024 threads takes jacobi_pipelined,
// which needs rc and rs of 64 entries.  Returns the number of sweeps.
__device__ inline int parallel_jacobi(double* A, double* V, int N, int ld, int* rp, int* rq, double* rc, double* rs,
                                      double* red) {
  if (N <= 64 && blockDim.x == 1024) return jacobi_pipelined(A, V, N, ld, rc, rs, red);
  const int m = N + (N & 1);  // even number of players; index N is a dummy
  const int half = m / 2;
  int sweep = 0;
  for (; sweep < MAX_SWEEPS; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int x = threadIdx.x; x < N * N; x += blockDim.x) {
      const int i = x / N, j = x - i * N;
      const double v = A[i * ld + j] * A[i * ld + j];
      tot += v;
      if (i != j) off += v;
    }
    off = block_sum(off, red);
    tot = block_sum(tot, red);
    // relative off-diagonal norm 1e-9: eigenvectors to ~1e-9, so D to ~1e-9 and
    // the energy (second order in D) to ~1e-18 of its scale
    if (off <= 1e-18 * tot) break;
    for (int r = 0; r < m - 1; ++r) {
      if ((int)threadIdx.x < half) {
        const int k = threadIdx.x;
        int p, q;
        if (k == 0) {
          p = m - 1;
          q = r;
        } else {  // (r + k) mod (m - 1), (r - k) mod (m - 1) without a division
          p = r + k;
          if (p >= m - 1) p -= m - 1;
          q = r - k;
          if (q < 0) q += m - 1;
        }
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double c = 1.0, s = 0.0;
        if (q < N) {
          const double apq = A[p * ld + q];
          const double app = A[p * ld + p], aqq = A[q * ld + q];
          if (fabs(apq) > 1e-300 && apq * apq > 1e-36 * fabs(app * aqq)) {
            // t = sign(th) / (|th| + sqrt(th^2 + 1)) with th = d / e, d = aqq - app,
            // e = 2 apq, written with one division: t = sign(d) e / (|d| + sqrt(d^2 + e^2))
            const double d = aqq - app, e = 2.0 * apq;
            const double t = (d >= 0 ? e : -e) / (fabs(d) + sqrt(d * d + e * e));
            c = rsqrt_nr(1.0 + t * t);
            s = t * c;
          }
        }
        rp[k] = p;
        rq[k] = q;
        rc[k] = c;
        rs[k] = s;
      }
      __syncthreads();
      // 2x2 blocks of A: block (k, l) = (tid / 32, tid % 32) when N <= 64
      const bool fast = half <= 32 && blockDim.x >= 1024;
      for (int b = threadIdx.x; b < (fast ? 1024 : half * half); b += blockDim.x) {
        const int k = fast ? (b >> 5) : b / half, l = fast ? (b & 31) : b - (b / half) * half;
        if (k >= half || l >= half) continue;
        const int pk = rp[k], qk = rq[k], pl = rp[l], ql = rq[l];
        const double ck = rc[k], sk = rs[k], cl = rc[l], sl = rs[l];
        const bool qkv = qk < N, qlv = ql < N;
        const double x00 = A[pk * ld + pl];
        const double x01 = qlv ? A[pk * ld + ql] : 0.0;
        const double x10 = qkv ? A[qk * ld + pl] : 0.0;
        const double x11 = (qkv && qlv) ? A[qk * ld + ql] : 0.0;
        // columns: [x.0, x.1] R_l
        const double t00 = cl * x00 - sl * x01, t01 = sl * x00 + cl * x01;
        const double t10 = cl * x10 - sl * x11, t11 = sl * x10 + cl * x11;
        // rows: R_k^T [t0.; t1.]
        A[pk * ld + pl] = ck * t00 - sk * t10;
        if (qlv) A[pk * ld + ql] = ck * t01 - sk * t11;
        if (qkv) A[qk * ld + pl] = sk * t00 + ck * t10;
        if (qkv && qlv) A[qk * ld + ql] = sk * t01 + ck * t11;
      }
      // columns of V (lanes walk rows i: stride ld, distinct banks); with
      // N <= 64, lane = row and wave = rotation (mod 16)
      for (int x = threadIdx.x; x < (fast ? 64 * half : half * N); x += blockDim.x) {
        const int l = fast ? (x >> 6) : x / N, i = fast ? (x & 63) : x - (x / N) * N;
        if (i >= N) continue;
        const int p = rp[l], q = rq[l];
        if (q >= N) continue;
        const double c = rc[l], s = rs[l];
        const double vip = V[i * ld + p], viq = V[i * ld + q];
        V[i * ld + p] = c * vip - s * viq;
        V[i * ld + q] = s * vip + c * viq;
      }
      __syncthreads();
    }
  }
  return sweep;
}

struct ScfLds {
  double *A, *U, *B, *Bst, *Bm, *Bv, *cvec, *red, *rc, *rs, *eig;
  int *rp, *rq, *rank;
};

constexpr int SCF_RED = (SCF_THREADS / 64) * DIIS_N;

// rotation tables hold N/2 + 1 pairs, at least 8 for the 10 x 10 DIIS system,
// and two rounds of 32 (jacobi_pipelined) for N <= 64
__host__ __device__ inline int scf_pairs(int N) { return N <= 64 ? 64 : N / 2 + 1; }

__host__ __device__ inline size_t scf_lds_bytes(int N) {
  const int h = scf_pairs(N), ld = lds_ld(N);
  return sizeof(double) * (3 * (size_t)N * ld + 300 + 16 + SCF_RED + 2 * h + N) + sizeof(int) * (2 * h + N);
}

__device__ inline ScfLds scf_lds_carve(unsigned char* smem, int N) {
  ScfLds L;
  const int NL = N * lds_ld(N), h = scf_pairs(N);
  L.A = reinterpret_cast<double*>(smem);
  L.U = L.A + NL;
  L.B = L.U + NL;
  L.Bst = L.B + NL;
  L.Bm = L.Bst + 100;
  L.Bv = L.Bm + 100;
  L.cvec = L.Bv + 100;
  L.red = L.cvec + 16;
  L.rc = L.red + SCF_RED;
  L.rs = L.rc + h;
  L.eig = L.rs + h;
  L.rp = reinterpret_cast<int*>(L.eig + N);
  L.rq = L.rp + h;
  L.rank = L.rq + h;
  return L;
}

// From the symmetric matrix in the orthonormal basis (L.A) and the initial
// rotation (L.U, identity or a warm start): Jacobi (eigenvectors optionally
// stored to Uout), ascending ranks of the eigenvalues, C = L^-T U (L^-1 in
// global Linv) left in L.A with columns unsorted, and C_occ C_occ^T (= D / 2)
// left in L.U (leading dimension ld).
__device__ inline int eig_density(const ScfLds& L, int N, int nocc, const double* __restrict__ Linv,
                                  double* __restrict__ Uout) {
  const int ld = lds_ld(N), tid = threadIdx.x;
  double *A = L.A, *U = L.U, *B = L.B;
  const int sweeps = parallel_jacobi(A, U, N, ld, L.rp, L.rq, L.rc, L.rs, L.red);
  if (Uout)  // eigenvectors in the orthonormal basis (next step's warm start)
    for (int x = tid; x < N * N; x += blockDim.x) {
      const int i = x / N;
      Uout[x] = U[i * ld + x - i * N];
    }
  for (int j = tid; j < N; j += blockDim.x) L.eig[j] = A[j * ld + j];
  lds_load(B, Linv, N, ld);
  __syncthreads();
  for (int j = tid; j < N; j += blockDim.x) {
    const double lj = L.eig[j];
    int r = 0;
    for (int k = 0; k < N; ++k) {
      const double lk = L.eig[k];
      r += (lk < lj) || (lk == lj && k < j);
    }
    L.rank[j] = r;
  }
  lds_matmul<true, false>(B, U, A, N, ld, ld);  // A = L^-T U (columns unsorted)
  __syncthreads();
  // occupied columns into B, then D = 2 B A^T into B via U
  for (int x = tid; x < N * N; x += blockDim.x) {
    const int i = x / N, j = x - i * N;
    B[i * ld + j] = L.rank[j] < nocc ? A[i * ld + j] : 0.0;
  }
  __syncthreads();
  lds_matmul<false, true>(B, A, U, N, ld, ld);  // U = C_occ C^T
  __syncthreads();
  return sweeps;
}

// One SCF step for one molecule by the whole workgroup (see pscf_scf_step).
// Optional Dt receives D converted to T (the precision of the next J/K / XC),
// optional dDt the change of D in T (for an incremental J/K build).
template <typename T, typename TJ>
__device__ inline void scf_step_body(const ScfLds& L, int it, int N, int nocc, const double* __restrict__ Hcore,
                                     const double* __restrict__ S, const double* __restrict__ Linv,
                                     const TJ* __restrict__ dJK, const T* __restrict__ Vxc, double* __restrict__ D,
                                     T* __restrict__ Dt, T* __restrict__ dDt, double* __restrict__ Hout,
                                     double* __restrict__ mo_e, double* __restrict__ Cout,
                                     double* __restrict__ state, int use_diis) {
  const int NN = N * N, ld = lds_ld(N);
  double *A = L.A, *U = L.U, *B = L.B;
  double* Vh = state;
  double* Hh = state + (size_t)DIIS_N * NN;
  double* Bst = Hh + (size_t)DIIS_N * NN;
  double* Uprev = Bst + 100;
  double* flags = Uprev + NN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = blockDim.x >> 6;

  // H = H_core + dJK + V_xc
  for (int x = tid; x < NN; x += blockDim.x) {
    const int i = x / N;
    A[i * ld + x - i * N] = Hcore[x] + (double)dJK[x] + (double)Vxc[x];
  }
  if (use_diis) {
    lds_load(B, D, N, ld);
    for (int x = tid; x < 100; x += blockDim.x) {
      const int i = x / 10, j = x % 10;
      L.Bst[x] = it == 0 ? (((i == 0) != (j == 0)) ? 1.0 : 0.0) : Bst[x];
    }
  }
  __syncthreads();

  if (use_diis) {
    const int head = it % DIIS_N;
    const int nvalid = it + 1 < DIIS_N ? it + 1 : DIIS_N;
    lds_matmul<false, false>(B, A, U, N, ld, ld);  // U = D H
    // H into the history (each thread reads back only its own elements below)
    for (int x = tid; x < NN; x += blockDim.x) {
      const int i = x / N;
      Hh[(size_t)head * NN + x] = A[i * ld + x - i * N];
    }
    __syncthreads();
    lds_load(B, S, N, ld);
    __syncthreads();
    lds_matmul<false, false>(B, U, A, N, ld, ld);  // A = S D H
    __syncthreads();
    // errvec = A - A^T into Vh[head]; <V_k, errvec> in one pass
    double part[DIIS_N];
#pragma unroll
    for (int k = 0; k < DIIS_N; ++k) part[k] = 0.0;
    double* err = Vh + (size_t)head * NN;
    // two elements per pass, all history loads issued before the error
    // vector's store (the store may alias those loads in the compiler's view)
    for (int x0 = tid; x0 < NN; x0 += 2 * blockDim.x) {
      double e[2], h[2][DIIS_N];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int x = x0 + u * blockDim.x;
        e[u] = 0.0;
        if (x < NN) {
          const int i = x / N, j = x - i * N;
          e[u] = A[i * ld + j] - A[j * ld + i];
        }
#pragma unroll
        for (int k = 0; k < DIIS_N; ++k)
          h[u][k] = (x < NN && k < nvalid && k != head) ? Vh[(size_t)k * NN + x] : e[u];
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
#pragma unroll
        for (int k = 0; k < DIIS_N; ++k)
          if (k < nvalid) part[k] += h[u][k] * e[u];
        const int x = x0 + u * blockDim.x;
        if (x < NN) err[x] = e[u];
      }
    }
#pragma unroll
    for (int k = 0; k < DIIS_N; ++k) {
      const double s = wave_sum(part[k]);
      if (lane == 0) L.red[wave * DIIS_N + k] = s;
    }
    __syncthreads();
    if (tid < nvalid) {
      double dot = 0.0;
      for (int w = 0; w < nwave; ++w) dot += L.red[w * DIIS_N + tid];
      L.Bst[(head + 1) * 10 + (tid + 1)] = dot;
      L.Bst[(tid + 1) * 10 + (head + 1)] = dot;
    }
    __syncthreads();
    if (wave == 0) {
      for (int x = lane; x < 100; x += 64) {
        const int i = x / 10, j = x % 10;
        const bool mi = i == 0 || i - 1 < nvalid, mj = j == 0 || j - 1 < nvalid;
        L.Bm[x] = (mi && mj) ? L.Bst[x] : 0.0;
      }
      wave_sync();
      wave_jacobi(L.Bm, L.Bv, 10, L.rp, L.rq, L.rc, L.rs);
      double wmax = 0.0;
      for (int i = 0; i < 10; ++i) wmax = fmax(wmax, fabs(L.Bm[i * 10 + i]));
      const double cut = 10.0 * 10.0 * 2.220446049250313e-16 * wmax;
      // pinv(B)[0, j] = sum_k V[0,k] V[j,k] / w_k over |w_k| > cut
      if (lane < 10) {
        double acc = 0.0;
        for (int k = 0; k < 10; ++k) {
          const double wk = L.Bm[k * 10 + k];
          if (fabs(wk) > cut) acc += L.Bv[k] * L.Bv[lane * 10 + k] / wk;
        }
        L.cvec[lane] = acc;
      }
    }
    __syncthreads();
    for (int x = tid; x < 100; x += blockDim.x) Bst[x] = L.Bst[x];
    for (int x = tid; x < NN; x += blockDim.x) {
      double h = 0.0;
      for (int k = 0; k < DIIS_N; ++k) h += L.cvec[k + 1] * Hh[(size_t)k * NN + x];
      const int i = x / N;
      A[i * ld + x - i * N] = h;
    }
  }
  lds_load(B, Linv, N, ld);
  __syncthreads();
  for (int x = tid; x < NN; x += blockDim.x) {
    const int i = x / N;
    Hout[x] = A[i * ld + x - i * N];
  }
  lds_matmul<false, true>(A, B, U, N, ld, ld);  // U = H L^T
  __syncthreads();
  lds_matmul<false, false>(B, U, A, N, ld, ld);  // A = L U
  __syncthreads();
  // exact symmetry (rounding of the FP32 Fock terms and of the transform
  // leaves O(eps) asymmetry that Jacobi rotations cannot remove)
  for (int x = tid; x < NN; x += blockDim.x) {
    const int i = x / N, j = x - i * N;
    if (i < j) {
      const double v = 0.5 * (A[i * ld + j] + A[j * ld + i]);
      A[i * ld + j] = v;
      A[j * ld + i] = v;
    }
  }
  __syncthreads();
  const bool warm = flags[0] != 0.0;
  if (warm) {
    lds_load(B, Uprev, N, ld);
    __syncthreads();
    lds_matmul<false, false>(A, B, U, N, ld, ld);  // U = A Uprev
    __syncthreads();
    lds_matmul<true, false>(B, U, A, N, ld, ld);  // A = Uprev^T U
    __syncthreads();
    for (int x = tid; x < N * ld; x += blockDim.x) U[x] = B[x];
  } else {
    for (int x = tid; x < N * ld; x += blockDim.x) U[x] = (x / ld == x % ld) ? 1.0 : 0.0;
  }
  __syncthreads();
  const int sweeps = eig_density(L, N, nocc, Linv, Uprev);
  if (tid == 0) {
    flags[0] = 1.0;
    flags[1] = (double)sweeps;  // diagnostic: Jacobi sweeps of this step
  }
  for (int j = tid; j < N; j += blockDim.x) mo_e[L.rank[j]] = L.eig[j];
  for (int x = tid; x < NN; x += blockDim.x) {
    const int i = x / N, j = x - i * N;
    Cout[i * N + L.rank[j]] = A[i * ld + j];
    const double d = 2.0 * U[i * ld + j];
    if (dDt) dDt[x] = (T)(d - D[x]);
    D[x] = d;
    if (Dt) Dt[x] = (T)d;
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
  const ScfLds L = scf_lds_carve(smem, N);
  scf_step_body<T, T>(L, it, N, nocc, Hcore, S, Linv, dJK, Vxc, D, (T*)nullptr, (T*)nullptr, Hout, mo_e, Cout, state,
                      use_diis);
}

// The same step for molecules whose three N x N working matrices exceed the
// LDS: they live in a caller workspace (L2-resident at N <= 512), the small
// DIIS and rotation tables stay in LDS.
constexpr int SCF_WS_MAXN = 512;

__host__ __device__ inline size_t scf_ws_doubles(int N) { return 3 * (size_t)N * lds_ld(N); }

template <typename T>
__global__ void __launch_bounds__(SCF_THREADS) scf_step_ws_kernel(int it, int N, int nocc, const double* __restrict__ Hcore,
                                                                  const double* __restrict__ S,
                                                                  const double* __restrict__ Linv,
                                                                  const T* __restrict__ dJK, const T* __restrict__ Vxc,
                                                                  double* __restrict__ D, double* __restrict__ Hout,
                                                                  double* __restrict__ mo_e, double* __restrict__ Cout,
                                                                  double* __restrict__ state, int use_diis,
                                                                  double* __restrict__ ws) {
  constexpr int H = SCF_WS_MAXN / 2 + 1;
  __shared__ double sm_d[300 + 16 + SCF_RED + 2 * H + SCF_WS_MAXN];
  __shared__ int sm_i[2 * H + SCF_WS_MAXN];
  ScfLds L;
  const int NL = N * lds_ld(N);
  L.A = ws;
  L.U = ws + NL;
  L.B = ws + 2 * (size_t)NL;
  L.Bst = sm_d;
  L.Bm = L.Bst + 100;
  L.Bv = L.Bm + 100;
  L.cvec = L.Bv + 100;
  L.red = L.cvec + 16;
  L.rc = L.red + SCF_RED;
  L.rs = L.rc + H;
  L.eig = L.rs + H;
  L.rp = sm_i;
  L.rq = L.rp + H;
  L.rank = L.rq + H;
  scf_step_body<T, T>(L, it, N, nocc, Hcore, S, Linv, dJK, Vxc, D, (T*)nullptr, (T*)nullptr, Hout, mo_e, Cout, state,
                      use_diis);
}

template <typename T>
__global__ void energy_kernel(int N, const double* __restrict__ D, const double* __restrict__ Hcore,
                              const T* __restrict__ dJK, const double* __restrict__ exc, double e_nuc,
                              double* __restrict__ out) {
  __shared__ double red[16];
  double s1 = 0.0, c1 = 0.0, s2 = 0.0, c2 = 0.0;  // Neumaier sums (1024 threads: a few terms each)
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

// ------------------------------------------------------------- batched ----
// Overlap Cholesky S = L L^T and L^-1 per molecule (FP64, LDS).  A molecule
// whose S is not positive definite is marked failed and inactive (ctl[2] = -1).
__global__ void __launch_bounds__(256) chol_batch_kernel(const BatchMol* __restrict__ mols,
                                                         const double* __restrict__ S, double* __restrict__ Linv,
                                                         double* __restrict__ ctl) {
  extern __shared__ __align__(16) unsigned char smem[];
  const BatchMol Mo = mols[blockIdx.x];
  const int N = (int)Mo.nao, NN = N * N;
  double* A = reinterpret_cast<double*>(smem);
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  for (int x = threadIdx.x; x < NN; x += blockDim.x) A[x] = S[Mo.mat0 + x];
  __syncthreads();
  for (int k = 0; k < N; ++k) {
    if (threadIdx.x == 0) {
      const double d = A[k * N + k];
      if (!(d > 0.0)) bad = 1;
      A[k * N + k] = d > 0.0 ? sqrt(d) : 1.0;
    }
    __syncthreads();
    const double lkk = A[k * N + k];
    for (int i = k + 1 + threadIdx.x; i < N; i += blockDim.x) A[i * N + k] /= lkk;
    __syncthreads();
    const int rem = N - k - 1;
    for (int t = threadIdx.x; t < rem * rem; t += blockDim.x) {
      const int i = k + 1 + t / rem, j = k + 1 + t % rem;
      if (j <= i) A[i * N + j] -= A[i * N + k] * A[j * N + k];
    }
    __syncthreads();
  }
  // L^-1 column by column (forward substitution), one thread per column
  double* out = Linv + Mo.mat0;
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    for (int i = 0; i < N; ++i) {
      if (i < j) {
        out[i * N + j] = 0.0;
        continue;
      }
      double acc = (i == j) ? 1.0 : 0.0;
      for (int k = j; k < i; ++k) acc -= A[i * N + k] * out[k * N + j];
      out[i * N + j] = acc / A[i * N + i];
    }
  }
  if (threadIdx.x == 0 && bad) {
    ctl[blockIdx.x * CTL + 1] = 0.0;
    ctl[blockIdx.x * CTL + 2] = -1.0;
  }
}

// Atomic-density guess (same rule as the host `guess.minao`): per atom and
// angular momentum, Gram-Schmidt the atom's shells (Cholesky of the S block)
// and occupy them with the element's spherical-average occupations.
__constant__ double c_occ[11][2][4] = {
    // {s occupations}, {p occupations (per component)} for Z = 0..10
    {{0, 0, 0, 0}, {0, 0, 0, 0}},
    {{1, 0, 0, 0}, {0, 0, 0, 0}},           // H  1s1
    {{2, 0, 0, 0}, {0, 0, 0, 0}},           // He
    {{2, 1, 0, 0}, {0, 0, 0, 0}},           // Li
    {{2, 2, 0, 0}, {0, 0, 0, 0}},           // Be
    {{2, 2, 0, 0}, {1.0 / 3, 0, 0, 0}},     // B
    {{2, 2, 0, 0}, {2.0 / 3, 0, 0, 0}},     // C
    {{2, 2, 0, 0}, {1.0, 0, 0, 0}},         // N
    {{2, 2, 0, 0}, {4.0 / 3, 0, 0, 0}},     // O
    {{2, 2, 0, 0}, {5.0 / 3, 0, 0, 0}},     // F
    {{2, 2, 0, 0}, {2.0, 0, 0, 0}}};        // Ne

template <typename T>
__global__ void __launch_bounds__(256) guess_batch_kernel(const BatchMol* __restrict__ mols,
                                                          const int* __restrict__ atm, const int* __restrict__ bas,
                                                          const int* __restrict__ shell_ao,
                                                          const double* __restrict__ S, double* __restrict__ D,
                                                          T* __restrict__ Dt) {
  const BatchMol Mo = mols[blockIdx.x];
  const int N = (int)Mo.nao, NN = N * N;
  double* Dm = D + Mo.mat0;
  const double* Sm = S + Mo.mat0;
  for (int x = threadIdx.x; x < NN; x += blockDim.x) Dm[x] = 0.0;
  __syncthreads();
  for (int ia = threadIdx.x; ia < (int)Mo.natm; ia += blockDim.x) {
    const int a = (int)Mo.atm0 + ia;
    const int z = atm[a * ATM_SLOTS];
    for (int l = 0; l < 2; ++l) {
      int ao[4], n = 0;
      for (int ib = (int)Mo.bas0; ib < (int)(Mo.bas0 + Mo.nbas) && n < 4; ++ib)
        if (bas[ib * BAS_SLOTS] == a && bas[ib * BAS_SLOTS + 1] == l) ao[n++] = shell_ao[ib];
      if (n == 0) continue;
      for (int m = 0; m < 2 * l + 1; ++m) {
        // S block -> L (Cholesky) -> C = L^-T (columns orthonormal)
        double Lb[4][4] = {{0}}, Ci[4][4] = {{0}};
        for (int i = 0; i < n; ++i)
          for (int j = 0; j <= i; ++j) {
            double acc = Sm[(ao[i] + m) * N + ao[j] + m];
            for (int k = 0; k < j; ++k) acc -= Lb[i][k] * Lb[j][k];
            Lb[i][j] = (i == j) ? sqrt(fmax(acc, 1e-300)) : acc / Lb[j][j];
          }
        for (int j = 0; j < n; ++j)  // Linv column j
          for (int i = j; i < n; ++i) {
            double acc = (i == j) ? 1.0 : 0.0;
            for (int k = j; k < i; ++k) acc -= Lb[i][k] * Ci[k][j];
            Ci[i][j] = acc / Lb[i][i];
          }
        // orbital k has coefficients C[:, k] = (L^-T)[:, k] = Linv[k][:]
        for (int k = 0; k < n; ++k) {
          const double o = c_occ[z > 10 ? 0 : z][l][k];
          if (o == 0.0) continue;
          for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) Dm[(ao[i] + m) * N + ao[j] + m] += o * Ci[k][i] * Ci[k][j];
        }
      }
    }
  }
  __syncthreads();
  if (Dt != nullptr)
    for (int x = threadIdx.x; x < NN; x += blockDim.x) Dt[Mo.mat0 + x] = (T)Dm[x];
}

template <typename T>
__global__ void __launch_bounds__(SCF_THREADS) scf_step_batch_kernel(const BatchMol* __restrict__ mols,
                                                                     const int* __restrict__ order,
                                                                     double* __restrict__ ctl,
                                                                     const double* __restrict__ Hcore,
                                                                     const double* __restrict__ S,
                                                                     const double* __restrict__ Linv,
                                                                     const double* __restrict__ dJK,
                                                                     const T* __restrict__ Vxc, double* __restrict__ D,
                                                                     T* __restrict__ Dt, T* __restrict__ dDt,
                                                                     double* __restrict__ H,
                                                                     double* __restrict__ mo_e,
                                                                     double* __restrict__ C, double* __restrict__ state,
                                                                     int use_diis) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int m = order[blockIdx.x];
  if (ctl[m * CTL + 1] == 0.0) return;
  const BatchMol Mo = mols[m];
  const int N = (int)Mo.nao;
  const ScfLds L = scf_lds_carve(smem, N);
  const int it = (int)ctl[m * CTL + 0];
  const int64_t o = Mo.mat0;
  scf_step_body<T, double>(L, it, N, (int)Mo.nocc, Hcore + o, S + o, Linv + o, dJK + o, Vxc + o, D + o, Dt + o,
                           dDt ? dDt + o : nullptr, H + o, mo_e + Mo.vec0, C + o, state + Mo.state0, use_diis);
}

// Purification of a predicted density (one workgroup per molecule):
// symmetrise, map to the orthonormal basis P = L^T D L = L^-1 (S D S) L^-T,
// diagonalise, and keep the nocc eigenvectors of largest occupation:
// D = 2 L^-T V_occ V_occ^T L^-1, so tr(D S) = 2 nocc and D S D = 2 D exactly.
template <typename T>
__global__ void __launch_bounds__(SCF_THREADS) purify_batch_kernel(const BatchMol* __restrict__ mols,
                                                                   const int* __restrict__ order,
                                                                   const double* __restrict__ ctl,
                                                                   const double* __restrict__ S,
                                                                   const double* __restrict__ Linv,
                                                                   double* __restrict__ D, T* __restrict__ Dt) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int m = order[blockIdx.x];
  if (ctl[m * CTL + 2] < 0.0) return;  // failed set-up
  const BatchMol Mo = mols[m];
  const int N = (int)Mo.nao, NN = N * N, nocc = (int)Mo.nocc, ld = lds_ld(N);
  const ScfLds L = scf_lds_carve(smem, N);
  double *A = L.A, *U = L.U, *B = L.B;
  const int64_t o = Mo.mat0;
  const int tid = threadIdx.x;
  for (int x = tid; x < NN; x += blockDim.x) {
    const int i = x / N, j = x - i * N;
    A[i * ld + j] = 0.5 * (D[o + x] + D[o + j * N + i]);
  }
  lds_load(B, S + o, N, ld);
  __syncthreads();
  lds_matmul<false, false>(B, A, U, N, ld, ld);  // U = S D
  __syncthreads();
  lds_matmul<false, false>(U, B, A, N, ld, ld);  // A = S D S
  __syncthreads();
  lds_load(B, Linv + o, N, ld);
  __syncthreads();
  lds_matmul<false, false>(B, A, U, N, ld, ld);  // U = L^-1 S D S
  __syncthreads();
  lds_matmul<false, true>(U, B, A, N, ld, ld);  // A = U L^-T
  __syncthreads();
  // -P, exactly symmetric: ascending eigenvalues = descending occupations
  for (int x = tid; x < NN; x += blockDim.x) {
    const int i = x / N, j = x - i * N;
    if (i <= j) {
      const double v = -0.5 * (A[i * ld + j] + A[j * ld + i]);
      A[i * ld + j] = v;
      A[j * ld + i] = v;
    }
  }
  for (int x = tid; x < N * ld; x += blockDim.x) U[x] = (x / ld == x % ld) ? 1.0 : 0.0;
  __syncthreads();
  eig_density(L, N, nocc, Linv + o, nullptr);
  for (int x = tid; x < NN; x += blockDim.x) {
    const int i = x / N, j = x - i * N;
    const double d = 2.0 * U[i * ld + j];
    D[o + x] = d;
    Dt[o + x] = (T)d;
  }
}

// Energy terms and convergence control per molecule.  init != 0: record E as
// E_prev only.  Otherwise: E_last = E, converged when |E - E_prev| < tol;
// iteration += 1; inactive once converged or at max_its.
template <typename T>
__global__ void __launch_bounds__(256) energy_batch_kernel(const BatchMol* __restrict__ mols,
                                                           const double* __restrict__ e_nuc,
                                                           const double* __restrict__ D,
                                                           const double* __restrict__ Hcore,
                                                           const double* __restrict__ dJK,
                                                           const double* __restrict__ exc,
                                                           const double* __restrict__ mo_e, double* __restrict__ ctl,
                                                           double* __restrict__ E, double* __restrict__ hist, int init,
                                                           double tol, int max_its) {
  __shared__ double red[16];
  const int m = blockIdx.x;
  if (!init && ctl[m * CTL + 1] == 0.0) return;
  if (init && ctl[m * CTL + 2] < 0.0) return;  // failed set-up
  const BatchMol Mo = mols[m];
  const int N = (int)Mo.nao;
  const int64_t o = Mo.mat0;
  double s1 = 0.0, c1 = 0.0, s2 = 0.0, c2 = 0.0;
  for (int x = threadIdx.x; x < N * N; x += blockDim.x) {
    const double a = D[o + x] * Hcore[o + x];
    double t = s1 + a;
    c1 += fabs(s1) >= fabs(a) ? (s1 - t) + a : (a - t) + s1;
    s1 = t;
    const double b = D[o + x] * (double)dJK[o + x];
    t = s2 + b;
    c2 += fabs(s2) >= fabs(b) ? (s2 - t) + b : (b - t) + s2;
    s2 = t;
  }
  const double e_core = block_sum(s1 + c1, red);
  const double e_jk = 0.5 * block_sum(s2 + c2, red);
  if (threadIdx.x == 0) {
    const double ex = exc[m];
    const double tot = e_core + e_jk + ex + e_nuc[m];
    double* Em = E + (int64_t)m * 5;
    Em[0] = tot;
    Em[1] = e_core;
    Em[2] = e_jk;
    Em[3] = ex;
    Em[4] = e_nuc[m];
    double* c = ctl + m * CTL;
    // ring of the last BATCH_HIST iteration energies (slot = iteration mod
    // BATCH_HIST; the initial energy is iteration 0; an init call on a
    // molecule that has already iterated leaves its ring alone)
    if (hist != nullptr && (!init || c[0] == 0.0)) hist[(int64_t)m * BATCH_HIST + ((init ? 0 : (int)c[0] + 1) % BATCH_HIST)] = tot;
    if (init) {
      c[3] = tot;
      c[4] = tot;
    } else {
      const int nocc = (int)Mo.nocc;
      if (nocc > 0 && nocc < N) c[5] = mo_e[Mo.vec0 + nocc] - mo_e[Mo.vec0 + nocc - 1];
      const double prev = c[3];
      c[4] = tot;
      c[3] = tot;
      c[0] += 1.0;
      c[6] = c[0];
      if (fabs(tot - prev) < tol) {
        c[2] = 1.0;
        c[1] = 0.0;
      } else if ((int)c[0] >= max_its) {
        c[1] = 0.0;
      }
    }
  }
}

// LDS for every molecule of nao <= N (a batch launch carves by each molecule's nao)
static size_t scf_lds(int N) {
  const size_t a = scf_lds_bytes(N), b = scf_lds_bytes(N < 64 ? N : 64);
  return a > b ? a : b;
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
  const int maxlds = device_max_lds();
  PSCF_REQUIRE((int)lds <= maxlds, "pscf_scf_step: N=%d needs %zu B LDS (max %d)", N, lds, maxlds);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64) {
    PSCF_SET_LDS(scf_step_kernel<double>, lds);
    scf_step_kernel<double><<<1, SCF_THREADS, lds, st>>>(iteration, N, nocc, d_Hcore, d_S, d_Linv,
                                                         (const double*)d_dJK, (const double*)d_Vxc, d_D, d_H,
                                                         d_mo_energy, d_C, (double*)d_state, use_diis);
  } else {
    PSCF_SET_LDS(scf_step_kernel<float>, lds);
    scf_step_kernel<float><<<1, SCF_THREADS, lds, st>>>(iteration, N, nocc, d_Hcore, d_S, d_Linv, (const float*)d_dJK,
                                                        (const float*)d_Vxc, d_D, d_H, d_mo_energy, d_C,
                                                        (double*)d_state, use_diis);
  }
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" size_t pscf_scf_step_ws_bytes(int N) { return sizeof(double) * scf_ws_doubles(N); }

extern "C" int pscf_scf_step_ws(int dtype, int iteration, int N, int nocc, const double* d_Hcore, const double* d_S,
                                const double* d_Linv, const void* d_dJK, const void* d_Vxc, double* d_D, double* d_H,
                                double* d_mo_energy, double* d_C, void* d_state, int use_diis, void* d_ws,
                                size_t ws_bytes, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_scf_step_ws: bad dtype");
  PSCF_REQUIRE(N > 1 && N <= SCF_WS_MAXN && nocc >= 0 && nocc <= N && iteration >= 0,
               "pscf_scf_step_ws: bad sizes (N <= %d)", SCF_WS_MAXN);
  PSCF_REQUIRE(d_Hcore && d_S && d_Linv && d_dJK && d_Vxc && d_D && d_H && d_mo_energy && d_C && d_state && d_ws,
               "pscf_scf_step_ws: null pointer");
  PSCF_REQUIRE(ws_bytes >= sizeof(double) * scf_ws_doubles(N), "pscf_scf_step_ws: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64)
    scf_step_ws_kernel<double><<<1, SCF_THREADS, 0, st>>>(iteration, N, nocc, d_Hcore, d_S, d_Linv,
                                                          (const double*)d_dJK, (const double*)d_Vxc, d_D, d_H,
                                                          d_mo_energy, d_C, (double*)d_state, use_diis, (double*)d_ws);
  else
    scf_step_ws_kernel<float><<<1, SCF_THREADS, 0, st>>>(iteration, N, nocc, d_Hcore, d_S, d_Linv, (const float*)d_dJK,
                                                         (const float*)d_Vxc, d_D, d_H, d_mo_energy, d_C,
                                                         (double*)d_state, use_diis, (double*)d_ws);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_energy(int dtype, int N, const double* d_D, const double* d_Hcore, const void* d_dJK,
                           const double* d_exc, double e_nuc, double* d_out, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_energy: bad dtype");
  PSCF_REQUIRE(N > 0 && d_D && d_Hcore && d_dJK && d_exc && d_out, "pscf_energy: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64)
    energy_kernel<double><<<1, 1024, 0, st>>>(N, d_D, d_Hcore, (const double*)d_dJK, d_exc, e_nuc, d_out);
  else
    energy_kernel<float><<<1, 1024, 0, st>>>(N, d_D, d_Hcore, (const float*)d_dJK, d_exc, e_nuc, d_out);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_batch_chol(const void* d_mols, int nmol, int max_nao, const double* d_S, double* d_Linv,
                               double* d_ctl, void* stream) {
  PSCF_REQUIRE(nmol > 0 && max_nao > 0 && d_mols && d_S && d_Linv && d_ctl, "pscf_batch_chol: bad arguments");
  const size_t lds = sizeof(double) * (size_t)max_nao * max_nao;
  const int maxlds = device_max_lds();
  PSCF_REQUIRE((int)lds <= maxlds, "pscf_batch_chol: nao %d too large", max_nao);
  PSCF_SET_LDS(chol_batch_kernel, lds);
  chol_batch_kernel<<<nmol, 256, lds, (hipStream_t)stream>>>((const BatchMol*)d_mols, d_S, d_Linv, d_ctl);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_batch_guess(int dtype, const void* d_mols, int nmol, const int* d_atm, const int* d_bas,
                                const int* d_shell_ao, const double* d_S, double* d_D, void* d_Dt, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_batch_guess: bad dtype");
  PSCF_REQUIRE(nmol > 0 && d_mols && d_atm && d_bas && d_shell_ao && d_S && d_D, "pscf_batch_guess: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64)
    guess_batch_kernel<double><<<nmol, 256, 0, st>>>((const BatchMol*)d_mols, d_atm, d_bas, d_shell_ao, d_S, d_D,
                                                     (double*)d_Dt);
  else
    guess_batch_kernel<float><<<nmol, 256, 0, st>>>((const BatchMol*)d_mols, d_atm, d_bas, d_shell_ao, d_S, d_D,
                                                    (float*)d_Dt);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_batch_scf_step(int dtype, const void* d_mols, const int* d_order, int nmol, int max_nao,
                                   double* d_ctl, const double* d_Hcore, const double* d_S, const double* d_Linv,
                                   const double* d_dJK, const void* d_Vxc, double* d_D, void* d_Dt, void* d_dDt,
                                   double* d_H, double* d_mo_energy, double* d_C, double* d_state, int use_diis,
                                   void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_batch_scf_step: bad dtype");
  PSCF_REQUIRE(nmol > 0 && max_nao > 1 && d_mols && d_order && d_ctl && d_Hcore && d_S && d_Linv && d_dJK && d_Vxc &&
                   d_D && d_Dt && d_H && d_mo_energy && d_C && d_state,
               "pscf_batch_scf_step: bad arguments");
  const size_t lds = scf_lds(max_nao);
  const int maxlds = device_max_lds();
  PSCF_REQUIRE((int)lds <= maxlds, "pscf_batch_scf_step: nao %d needs %zu B LDS", max_nao, lds);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64) {
    PSCF_SET_LDS(scf_step_batch_kernel<double>, lds);
    scf_step_batch_kernel<double><<<nmol, SCF_THREADS, lds, st>>>(
        (const BatchMol*)d_mols, d_order, d_ctl, d_Hcore, d_S, d_Linv, d_dJK, (const double*)d_Vxc, d_D,
        (double*)d_Dt, (double*)d_dDt, d_H, d_mo_energy, d_C, d_state, use_diis);
  } else {
    PSCF_SET_LDS(scf_step_batch_kernel<float>, lds);
    scf_step_batch_kernel<float><<<nmol, SCF_THREADS, lds, st>>>(
        (const BatchMol*)d_mols, d_order, d_ctl, d_Hcore, d_S, d_Linv, d_dJK, (const float*)d_Vxc, d_D,
        (float*)d_Dt, (float*)d_dDt, d_H, d_mo_energy, d_C, d_state, use_diis);
  }
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_batch_purify(int dtype, const void* d_mols, const int* d_order, int nmol, int max_nao,
                                 const double* d_ctl, const double* d_S, const double* d_Linv, double* d_D, void* d_Dt,
                                 void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_batch_purify: bad dtype");
  PSCF_REQUIRE(nmol > 0 && max_nao > 1 && d_mols && d_order && d_ctl && d_S && d_Linv && d_D && d_Dt,
               "pscf_batch_purify: bad arguments");
  const size_t lds = scf_lds(max_nao);
  const int maxlds = device_max_lds();
  PSCF_REQUIRE((int)lds <= maxlds, "pscf_batch_purify: nao %d needs %zu B LDS", max_nao, lds);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64) {
    PSCF_SET_LDS(purify_batch_kernel<double>, lds);
    purify_batch_kernel<double><<<nmol, SCF_THREADS, lds, st>>>((const BatchMol*)d_mols, d_order, d_ctl, d_S, d_Linv,
                                                                d_D, (double*)d_Dt);
  } else {
    PSCF_SET_LDS(purify_batch_kernel<float>, lds);
    purify_batch_kernel<float><<<nmol, SCF_THREADS, lds, st>>>((const BatchMol*)d_mols, d_order, d_ctl, d_S, d_Linv,
                                                               d_D, (float*)d_Dt);
  }
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_batch_energy(int dtype, const void* d_mols, int nmol, const double* d_e_nuc, const double* d_D,
                                 const double* d_Hcore, const double* d_dJK, const double* d_exc,
                                 const double* d_mo_energy, double* d_ctl, double* d_E, double* d_hist, int init,
                                 double tol, int max_its, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_batch_energy: bad dtype");
  PSCF_REQUIRE(nmol > 0 && d_mols && d_e_nuc && d_D && d_Hcore && d_dJK && d_exc && d_mo_energy && d_ctl && d_E,
               "pscf_batch_energy: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64)
    energy_batch_kernel<double><<<nmol, 256, 0, st>>>((const BatchMol*)d_mols, d_e_nuc, d_D, d_Hcore, d_dJK, d_exc,
                                                      d_mo_energy, d_ctl, d_E, d_hist, init, tol, max_its);
  else
    energy_batch_kernel<float><<<nmol, 256, 0, st>>>((const BatchMol*)d_mols, d_e_nuc, d_D, d_Hcore, d_dJK, d_exc,
                                                     d_mo_energy, d_ctl, d_E, d_hist, init, tol, max_its);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}
