// SCF step (DIIS + generalized eigensolve + density) and energy on gfx950.
//
// Reference: nanoDFT_iteration (pyscf_ipu/nanoDFT/nanoDFT.py:40-102): H = H_core
// + dJK + V_xc; DIIS (:338-364): errvec = S D H - (S D H)^T into a 9-slot ring,
// B[head+1, 1:] = <V_k, errvec>, c = pinv(masked B)[0], H = sum c_k H_k;
// eigvects = L_inv^T eigh(L_inv H L_inv^T)[1] (:69, linalg_eigh :370-393);
// D = 2 C_occ C_occ^T (:72); energy (:20-38).
// One workgroup does the whole step in FP64 with the working matrices in LDS.
// The eigensolver (N <= 64) is a tridiagonal solver on the first step and
// wherever the previous step's eigenvectors leave A' = U_prev^T A U_prev far
// from diagonal, and otherwise an iterative refinement of U_prev on MFMA
// products (Ogita-Aishima); a parallel cyclic Jacobi (round-robin pairing, N/2
// disjoint rotations per round) is the fallback and serves N > 64.
#include "batch.h"
#include "common.h"
#include "wave.h"

namespace pscf {

constexpr int SCF_THREADS = 1024;
constexpr int DIIS_N = 9;
constexpr int MAX_SWEEPS = 30;

// state layout (doubles): Vhist[9][N*N] | Hhist[9][N*N] | B[10][10] | Wprev[N*N] | flags[2]
// (Wprev: the previous eigenvectors, C = L^-T U for N <= 64 with 1024 threads, else U)
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

// The same product with every operand fragment of a tile loaded before its
// MFMA chain (eight k-steps at a time), so that LDS latency overlaps the
// matrix unit (64 cycles per FP64 16x16x4 operation per SIMD): about 6k
// cycles for N = 64 against 8.5k loading per step.  kSym: C is symmetric,
// only tiles I <= J are computed and each is also stored transposed.
template <bool TX, bool TY, bool kSym>
__device__ inline void lds_matmul_mfma_pre(const double* __restrict__ X, const double* __restrict__ Y,
                                           double* __restrict__ C, int N, int ld, int ldc) {
  typedef double v4d __attribute__((ext_vector_type(4)));
  const int nt = (N + 15) >> 4;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int I, J;
  if (kSym) {  // wave w -> the w-th upper tile (I <= J), row by row
    int t = w, r = 0;
    while (r < nt && t >= nt - r) {
      t -= nt - r;
      ++r;
    }
    if (r >= nt) return;  // wave-uniform
    I = r;
    J = r + t;
  } else {
    if (w >= nt * nt) return;  // wave-uniform
    I = w / nt;
    J = w - I * nt;
  }
  const int li = lane & 15, lg = lane >> 4;
  const int row = I * 16 + li, col = J * 16 + li;
  const bool rv = row < N, cv = col < N;
  const int rowc = rv ? row : N - 1, colc = cv ? col : N - 1;
  v4d acc = {0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < N; k0 += 32) {
    double a[8], b[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int k = k0 + 4 * t + lg;
      const bool kv = k < N;
      const int kc = kv ? k : N - 1;
      const double av = TX ? X[kc * ld + rowc] : X[rowc * ld + kc];
      const double bv = TY ? Y[colc * ld + kc] : Y[kc * ld + colc];
      a[t] = (rv && kv) ? av : 0.0;
      b[t] = (cv && kv) ? bv : 0.0;
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b[t], acc, 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int r = I * 16 + lg + 4 * j;
    if (r < N && cv && (!kSym || I < J || r <= col)) {  // a diagonal tile stores its upper half
      C[r * ldc + col] = acc[j];
      if (kSym) C[col * ldc + r] = acc[j];
    }
  }
}

// C = op(X) op(Y), N x N, X and Y in LDS (leading dimension ld), C in LDS or
// global (leading dimension ldc), never aliasing X or Y.  N <= 64 with 1024
// threads takes the MFMA forms above.  Otherwise each thread owns rows
// {i, i + h} x columns {j, j + h} (h = ceil(N / 2)), so the lanes of a wave walk
// consecutive columns and share rows.
template <bool TX, bool TY>
__device__ inline void lds_matmul(const double* __restrict__ X, const double* __restrict__ Y, double* __restrict__ C,
                                  int N, int ld, int ldc) {
  if (N <= 64 && blockDim.x == 1024) {
    if (N > 32)
      lds_matmul_mfma_pre<TX, TY, false>(X, Y, C, N, ld, ldc);
    else
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

// lds_matmul for a product known to be symmetric (C = C^T up to rounding):
// for 32 < N <= 64 only the tiles on and above the diagonal are computed and
// mirrored, so C comes out exactly symmetric.
template <bool TX, bool TY>
__device__ inline void lds_matmul_sym(const double* __restrict__ X, const double* __restrict__ Y,
                                      double* __restrict__ C, int N, int ld, int ldc) {
  if (N > 32 && N <= 64 && blockDim.x == 1024)
    lds_matmul_mfma_pre<TX, TY, true>(X, Y, C, N, ld, ldc);
  else
    lds_matmul<TX, TY>(X, Y, C, N, ld, ldc);
}

// N x N global (leading dimension N) <-> LDS (leading dimension ld)
__device__ inline void lds_load(double* __restrict__ dst, const double* __restrict__ src, int N, int ld) {
  for (int x = threadIdx.x; x < N * N; x += blockDim.x) {
    const int i = x / N;
    dst[i * ld + x - i * N] = src[x];
  }
}

// Sum over the 16 lanes of each DPP row (lanes 16 r .. 16 r + 15), in every
// lane of the row: quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror,
// row_mirror (VALU moves, no LDS traffic).  The whole wavefront must be active.
__device__ inline double row16_sum(double v) {
  v = dpp_add<0xb1, 0xf>(v);
  v = dpp_add<0x4e, 0xf>(v);
  v = dpp_add<0x141, 0xf>(v);
  return dpp_add<0x140, 0xf>(v);
}

// Lane `src`'s value of v (src uniform across the wavefront).
__device__ inline double readlane_f64(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes, in every lane: row sums, then the rows added by the
// permlane16 / permlane32 swaps.  The whole wavefront must be active.
__device__ inline double wave_sum(double v) { return xor32_sum(xor16_sum(row16_sum(v))); }

template <int CTRL>
__device__ inline double dpp_max(double v) {
  return fmax(v, dpp_t<CTRL, 0xf>(v));
}

// Maximum over the 64 lanes, in every lane (the whole wavefront active).
__device__ inline double wave_max(double v) {
  v = dpp_max<0xb1>(v);
  v = dpp_max<0x4e>(v);
  v = dpp_max<0x141>(v);
  v = dpp_max<0x140>(v);
  const auto h = __builtin_amdgcn_permlane16_swap(__double2hiint(v), __double2hiint(v), false, false);
  const auto l = __builtin_amdgcn_permlane16_swap(__double2loint(v), __double2loint(v), false, false);
  v = fmax(__hiloint2double(h[0], l[0]), __hiloint2double(h[1], l[1]));
  const auto h2 = __builtin_amdgcn_permlane32_swap(__double2hiint(v), __double2hiint(v), false, false);
  const auto l2 = __builtin_amdgcn_permlane32_swap(__double2loint(v), __double2loint(v), false, false);
  return fmax(__hiloint2double(h2[0], l2[0]), __hiloint2double(h2[1], l2[1]));
}

// Sum (max) of the per-wavefront values part[0..n), n <= 16, in every lane of
// the calling (whole) wavefront: lanes 0..15 load one value each.
__device__ inline double combine_sum(const double* part, int n) {
  const int lane = threadIdx.x & 63;
  return readlane_f64(row16_sum(lane < n ? part[lane] : 0.0), 0);
}
__device__ inline double combine_max(const double* part, int n) {
  const int lane = threadIdx.x & 63;
  double v = lane < n ? part[lane] : -1.0 / 0.0;
  v = dpp_max<0xb1>(v);
  v = dpp_max<0x4e>(v);
  v = dpp_max<0x141>(v);
  return readlane_f64(dpp_max<0x140>(v), 0);
}

__device__ inline double block_sum(double v, double* red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  const double s = combine_sum(red, (int)(blockDim.x >> 6));
  __syncthreads();
  return s;
}

// Orders the LDS accesses of one wavefront (its lanes run in lockstep).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
__device__ __forceinline__ int jacobi_pipelined(double* A, double* V, int N, int ld, double* rc, double* rs, double* red,
                                       double stop) {
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
    if (off <= stop * tot) break;
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
// which needs rc and rs of 64 entries.  Sweeps until the off-diagonal share of
// |A|^2 is at most `stop`; returns the number of sweeps.
__device__ __forceinline__ int parallel_jacobi(double* A, double* V, int N, int ld, int* rp, int* rq, double* rc, double* rs,
                                      double* red, double stop = 1e-18) {
  if (N <= 64 && blockDim.x == 1024) return jacobi_pipelined(A, V, N, ld, rc, rs, red, stop);
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
    // stop = 1e-18, a relative off-diagonal norm of 1e-9: eigenvectors to ~1e-9,
    // so D to ~1e-9 and the energy (second order in D) to ~1e-18 of its scale
    if (off <= stop * tot) break;
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
  double* tri;  // scratch of tri_eigh (N <= 64 with 1024 threads), else null
  double* M;    // fourth N x ld matrix (N <= 64 with 1024 threads), else null
};

// tri_eigh scratch: d, e, lo, hi, w, lam, dot (64 doubles each), kpart (16);
// then ints: cnt[16][64], block lo / hi (64 each), fail flag
constexpr int TRI_DOUBLES = 7 * 64 + 16;
constexpr int TRI_INTS = 16 * 64 + 2 * 64 + 4;
constexpr size_t TRI_BYTES = sizeof(double) * TRI_DOUBLES + sizeof(int) * TRI_INTS;
// multisection: sections per round (one wavefront each, one per SIMD: the
// Sturm recurrence is bound by instruction issue) and rounds, 5^23 > 2^53: the
// intervals shrink to the rounding of their end points
constexpr int TRI_SECT = 4;
constexpr int TRI_ROUNDS = 23;

constexpr int SCF_RED = (SCF_THREADS / 64) * DIIS_N;

// rotation tables hold N/2 + 1 pairs, at least 8 for the 10 x 10 DIIS system,
// and two rounds of 32 (jacobi_pipelined) for N <= 64
__host__ __device__ inline int scf_pairs(int N) { return N <= 64 ? 64 : N / 2 + 1; }

__host__ __device__ inline size_t scf_lds_bytes(int N) {
  const int h = scf_pairs(N), ld = lds_ld(N);
  return sizeof(double) * (3 * (size_t)N * ld + 300 + 16 + SCF_RED + 2 * h + N) + sizeof(int) * (2 * h + N) +
         (N <= 64 ? 8 + TRI_BYTES + sizeof(double) * (size_t)N * ld : 0);
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
  L.tri = nullptr;
  L.M = nullptr;
  if (N <= 64) {  // 8-byte aligned after the int tables
    const size_t off = (reinterpret_cast<unsigned char*>(L.rank + N) - smem + 7) & ~(size_t)7;
    L.tri = reinterpret_cast<double*>(smem + off);
    L.M = reinterpret_cast<double*>(smem + off + TRI_BYTES);
  }
  return L;
}

// ------------------------------------------------- tridiagonal eigensolver ---
// Symmetric eigen-decomposition for N <= 64 by 1024 threads, in place of the
// Jacobi sweeps: a fixed number of barrier phases instead of a number of
// sweeps that depends on how far A is from diagonal.
//   1. Householder tridiagonalisation T = Q^T A Q with A and Q held in the
//      registers of 512 threads (8 per row): step k reflects column k below
//      the diagonal onto e_(k+1) (H = I - tau u u^T), p = tau A u, K = tau
//      u.p / 2, w = p - K u, A -= u w^T + w u^T, Q -= tau (Q u) u^T.
//   2. Off-diagonals below 1e-12 (|d_i| + |d_i+1|) are set to zero: T falls
//      into unreduced blocks, so eigenvalues that are degenerate up to
//      rounding come from different blocks.
//   3. Eigenvalue slot k belongs to the block holding position k (a block of
//      n rows owns n slots): bisection on the block's Sturm count, each of the
//      16 waves evaluating one of 16 interior points of every slot's interval
//      per round (width / 17 per round, 14 rounds from the Gershgorin
//      interval).  The count uses the division-free recurrence
//      p_i = (d_i - x) p_(i-1) - e_(i-1)^2 p_(i-2) with power-of-two rescaling
//      and counts sign changes (a zero takes the opposite sign of its
//      predecessor).
//   4. Eigenvector of slot k from the twisted factorisation of T - lambda I on
//      its block (Dp top-down, Dm bottom-up, gamma_i = Dp_i + Dm_i - (d_i -
//      lambda), r = argmin |gamma|, z_r = 1, z_i = -e_i / Dp_i z_(i+1) above,
//      z_i = -e_(i-1) / Dm_i z_(i-1) below), one lane per slot, up to 32
//      slots at a time (1 / Dp and 1 / Dm of a slot in B).
//   5. Columns in ascending eigenvalue order; eigenvalues closer than 1e-6 of
//      the scale are orthogonalised in order (modified Gram-Schmidt).
//   6. U = Q Z by applying the reflectors in reverse order.
// Returns false (the caller then runs Jacobi) if a Gram-Schmidt residual
// falls below 1/2 or a result is not finite.  On success U holds the
// eigenvectors as columns, eig the eigenvalues, both ascending.
__device__ inline double recip_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = y * (2.0 - x * y);
  return y * (2.0 - x * y);
}

// Sturm count of the block [b0, b1) of the tridiagonal matrix (diagonal d,
// e2m1[i] = e_(i-1)^2, zero where the matrix splits): the number of its
// eigenvalues below x.  The product recurrence p_i = (d_i - x) p_(i-1) -
// e_(i-1)^2 p_(i-2) runs over all N positions with a wavefront-uniform index
// (d_i, e2m1_i broadcast from LDS), counting sign changes inside [b0, b1)
// only: a zero coupling at the block start makes sign(p_b0 / p_(b0-1)) =
// sign(d_b0 - x), the block's own first term.  Branch-free (sign bits and
// masks), rescaled by a power of two every 8 positions; a zero counts with the
// sign of its sign bit.  d and e2m1 have 64 entries, zero from N on (the last
// chunk of 8 runs past N without counting).
__device__ inline int sturm_count_all(const double* __restrict__ d, const double* __restrict__ e2m1, int N, int b0,
                                      int b1, double x) {
  double p = 1.0, pm = 0.0;
  int sp = 0, c = 0;  // sign of the previous term: 0 non-negative, -1 negative
  const int hi = b1 - 1;
  for (int i0 = 0; i0 < N; i0 += 8) {  // positions N..63 hold d = 0, e2m1 = 0 and count nowhere
    double dd[8], ee[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      dd[u] = d[i0 + u] - x;
      ee[u] = e2m1[i0 + u];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u;
      const double pn = dd[u] * p - ee[u] * pm;
      pm = p;
      p = pn;
      const int sg = __double2hiint(pn) >> 31;
      c -= (sg ^ sp) & (~((i - b0) | (hi - i)) >> 31);
      sp = sg;
    }
    const double mx = fmax(fabs(p), fabs(pm));
    const int ex = (mx > 0x1p200 || mx < 0x1p-200) ? __builtin_amdgcn_frexp_exp(mx) : 0;
    p = ldexp(p, -ex);
    pm = ldexp(pm, -ex);
  }
  return c;
}

// sturm_count_all for a matrix without splits (every slot's block is [0, N)):
// the full chunks of 8 positions count every sign change, only the last
// partial chunk masks the positions from N on.
__device__ inline int sturm_count_whole(const double* __restrict__ d, const double* __restrict__ e2m1, int N,
                                        double x) {
  double p = 1.0, pm = 0.0;
  int sp = 0, c = 0;
  auto chunk = [&](int i0, bool tail) {
    double dd[8], ee[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      dd[u] = d[i0 + u] - x;
      ee[u] = e2m1[i0 + u];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const double pn = dd[u] * p - ee[u] * pm;
      pm = p;
      p = pn;
      const int sg = __double2hiint(pn) >> 31;
      c -= tail ? ((sg ^ sp) & ((i0 + u - N) >> 31)) : (sg ^ sp);
      sp = sg;
    }
    const double mx = fmax(fabs(p), fabs(pm));
    const int ex = (mx > 0x1p200 || mx < 0x1p-200) ? __builtin_amdgcn_frexp_exp(mx) : 0;
    p = ldexp(p, -ex);
    pm = ldexp(pm, -ex);
  };
  int i0 = 0;
  for (; i0 + 8 <= N; i0 += 8) chunk(i0, false);
  if (i0 < N) chunk(i0, true);
  return c;
}

// Eigenvectors of all slots (lane = slot, eigenvalue lk, column col of Z) of
// a tridiagonal matrix without splits, by the twisted factorisation of
// twisted_slot with every block [0, N).  Branch-free: the positions run in
// chunks of 8 whose d, e and pivot values are loaded before the recurrence
// over them (only the last chunk checks i < N), a pivot below tiny in
// magnitude becomes +-tiny by fmax / copysign, and stores of inactive lanes
// (or of positions a lane does not own) go to the one-double sink.  The
// vectors are stored unnormalised; scale[col] receives 1 / |z|.  tiny is eps
// max_i (|d_i| + |e_i|) + 1e-300.  d and e have 64 entries.
__device__ inline void twisted_whole(const double* __restrict__ d, const double* __restrict__ e, int N, int ld,
                                     double tiny, double lk, bool act, int col, double* __restrict__ rdp,
                                     double* __restrict__ rdm, double* __restrict__ Z, double* __restrict__ scale,
                                     double* __restrict__ sink) {
  auto guard = [&](double v) { return copysign(fmax(fabs(v), tiny), v); };
  double r_prev = 0.0;
  auto down = [&](int i0, bool tail) {  // top-down pivots of positions i0 .. i0 + 7
    double dd[8], em[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      dd[u] = d[i0 + u] - lk;
      em[u] = i0 + u > 0 ? e[i0 + u - 1] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (tail && i0 + u >= N) break;
      r_prev = recip_nr(guard(dd[u] - em[u] * em[u] * r_prev));
      *(act ? rdp + i0 + u : sink) = r_prev;
    }
  };
  int i0 = 0;
  for (; i0 + 8 <= N; i0 += 8) down(i0, false);
  if (i0 < N) down(i0, true);
  double rnext = 0.0, gbest = 1.0 / 0.0;
  int r = N - 1;
  auto up = [&](int i0, bool tail) {  // bottom-up pivots and the twist, positions i0, i0 - 1, .. i0 - 7
    double dd[8], ei[8], em[8], rp[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 - u;
      const int ic = i > 0 ? i : 0;
      dd[u] = d[ic] - lk;
      ei[u] = i < N - 1 ? e[ic] : 0.0;
      em[u] = i > 0 ? e[ic - 1] : 0.0;
      rp[u] = i > 0 ? rdp[ic - 1] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 - u;
      if (tail && i < 0) break;
      const double dm = guard(dd[u] - ei[u] * ei[u] * rnext);
      rnext = recip_nr(dm);
      *(act ? rdm + i : sink) = rnext;
      const double gm = fabs(i > 0 ? dm - em[u] * em[u] * rp[u] : dm);
      r = gm < gbest ? i : r;
      gbest = fmin(gm, gbest);
    }
  };
  i0 = N - 1;
  for (; i0 >= 7; i0 -= 8) up(i0, false);
  if (i0 >= 0) up(i0, true);
  *(act ? Z + r * ld + col : sink) = 1.0;
  double nrm = 1.0, z = 1.0;
  auto above = [&](int i0, bool tail) {  // z at positions i0, i0 - 1, .. below r
    double ee[8], rr[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int ic = i0 - u > 0 ? i0 - u : 0;
      ee[u] = e[ic];
      rr[u] = rdp[ic];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 - u;
      if (tail && i < 0) break;
      const bool own = act && i < r;
      const double zn = -ee[u] * rr[u] * z;
      z = own ? zn : z;
      nrm += own ? zn * zn : 0.0;
      *(own ? Z + i * ld + col : sink) = zn;
    }
  };
  i0 = N - 2;
  for (; i0 >= 7; i0 -= 8) above(i0, false);
  if (i0 >= 0) above(i0, true);
  z = 1.0;
  auto below = [&](int i0, bool tail) {  // z at positions i0, i0 + 1, .. above r
    double ee[8], rr[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int ic = i0 + u < N ? i0 + u : N - 1;
      ee[u] = e[ic - 1];
      rr[u] = rdm[ic];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u;
      if (tail && i >= N) break;
      const bool own = act && i > r;
      const double zn = -ee[u] * rr[u] * z;
      z = own ? zn : z;
      nrm += own ? zn * zn : 0.0;
      *(own ? Z + i * ld + col : sink) = zn;
    }
  };
  i0 = 1;
  for (; i0 + 8 <= N; i0 += 8) below(i0, false);
  if (i0 < N) below(i0, true);
  if (act) scale[col] = 1.0 / sqrt(nrm);
}

// Eigenvector of one slot (lane) of the tridiagonal matrix (d, e in LDS) for
// the eigenvalue lk of its block [a0, a1), by the twisted factorisation:
// top-down pivots Dp, bottom-up pivots Dm, the twist r where |gamma_i| = |Dm_i
// - e_(i-1)^2 / Dp_(i-1)| is smallest, then z_r = 1 and the recurrences up and
// down from r, normalised, into column col of Z.  Pivots smaller than eps
// |T| are replaced by +-eps |T|.  kPred: per-lane block bounds (the matrix
// splits); otherwise every slot's block is [0, N) and the positions run with
// one uniform index.  Inactive lanes (act false) store nothing.
template <bool kPred>
__device__ inline void twisted_slot(const double* __restrict__ d, const double* __restrict__ e, int N, int ld, int a0,
                                    int a1, double lk, bool act, int col, double* __restrict__ rdp,
                                    double* __restrict__ rdm, double* __restrict__ Z) {
  auto in = [&](int i) { return !kPred || (i >= a0 && i < a1); };
  double sc = 0.0;
  for (int i = 0; i < N; ++i)
    if (in(i)) sc = fmax(sc, fabs(d[i]) + (i < a1 - 1 ? fabs(e[i]) : 0.0));
  const double tiny = 2.220446049250313e-16 * sc + 1e-300;
  auto guard = [&](double v) { return fabs(v) >= tiny ? v : (v < 0.0 ? -tiny : tiny); };
  double r_prev = 0.0;
  for (int i = 0; i < N; ++i) {
    if (in(i)) {
      const double em = i > a0 ? e[i - 1] : 0.0;
      r_prev = recip_nr(guard((d[i] - lk) - em * em * r_prev));
      if (act) rdp[i] = r_prev;
    }
  }
  double rnext = 0.0, gbest = 1.0 / 0.0;
  int r = a1 - 1;
  for (int i = N - 1; i >= 0; --i) {
    if (in(i)) {
      const double ei = i < a1 - 1 ? e[i] : 0.0, em = i > a0 ? e[i - 1] : 0.0;
      const double dm = guard((d[i] - lk) - ei * ei * rnext);
      rnext = recip_nr(dm);
      if (act) rdm[i] = rnext;
      const double g = fabs(i > a0 ? dm - em * em * rdp[i - 1] : dm);
      if (g < gbest) {
        gbest = g;
        r = i;
      }
    }
  }
  if (!act) return;
  double z = 1.0, nrm = 1.0;
  Z[r * ld + col] = 1.0;
  for (int i = r - 1; i >= a0; --i) {
    z = -e[i] * rdp[i] * z;
    Z[i * ld + col] = z;
    nrm += z * z;
  }
  z = 1.0;
  for (int i = r + 1; i < a1; ++i) {
    z = -e[i - 1] * rdm[i] * z;
    Z[i * ld + col] = z;
    nrm += z * z;
  }
  const double inv = 1.0 / sqrt(nrm);
  for (int i = a0; i < a1; ++i) Z[i * ld + col] *= inv;
}

// Buffers of one tri_eigh call: A (input, overwritten) and Z (eigenvectors)
// with leading dimension ld, W scratch of wcap doubles (at least N ld), eig /
// rank (N), the solver's own scratch region `tri` (TRI_BYTES), optionally W2,
// a second N x ld scratch matrix (with it the eigenvectors of all slots are
// formed in one pass, from register-held d and e), and optionally Q0, an
// orthogonal N x N matrix in global memory (leading dimension N): Z then
// receives Q0 times the eigenvectors of A (those of Q0 A Q0^T).
struct TriBufs {
  double *A, *Z, *W, *eig;
  int* rank;
  double* tri;
  int ld, wcap;
  double* W2;
  const double* Q0;
};

__device__ inline bool tri_eigh(const TriBufs& T, int N) {
  const int ld = T.ld, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double *A = T.A, *Z = T.Z, *W = T.W;
  double* d = T.tri;
  double* e = d + 64;
  double* lo = e + 64;
  double* hi = lo + 64;
  double* wv = hi + 64;
  double* lam = wv + 64;
  double* dot = lam + 64;
  double* kpart = dot + 64;
  int* cnt = reinterpret_cast<int*>(kpart + 16);
  int* blo = cnt + 16 * 64;
  int* bhi = blo + 64;
  int* fail = bhi + 64;
  // 1. tridiagonalisation, register-resident in the first 512 threads (two
  // wavefronts per SIMD): thread (row ri = tid / 8, g = tid % 8) holds A[ri][c]
  // and Q[ri][c] for its eight columns c = g + 8 u (zero outside N x N), Q =
  // Q0 (or the identity) to start.  Step k reflects column k below the
  // diagonal, x = A[k+1..N-1][k], onto e_(k+1): alpha = -sign(x_0) |x|, u = x
  // - alpha e_(k+1) (zero at positions <= k), H = I - tau u u^T with tau =
  // 2 / u^T u = 1 / (|x| (|x| + |x_0|)).  Then p = tau A u, K = tau (u^T p) / 2,
  // w = p - K u, A <- A - u w^T - w u^T (the two-sided H A H, applied to the
  // whole matrix) and Q <- Q - tau (Q u) u^T, so that Q = Q0 H_0 H_1 ... needs
  // no back-transformation pass.  The 8 lanes of a row hold all 64 positions
  // between them, so row sums (DPP, no LDS) give A u, Q u, u^T p and |x|^2.
  // Two barriers a step: p into LDS; the updated column k + 1 into LDS, from
  // which every wavefront forms the next reflector itself (identical
  // operations on identical values in every row, so every row holds the same
  // u and tau).
  const bool red = tid < 512;
  const int ri = (tid >> 3) & 63, g = tid & 7;
  double ar[8], qr[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int c = g + 8 * u;
    const bool in = red && ri < N && c < N;
    ar[u] = in ? A[ri * ld + c] : 0.0;
    qr[u] = in ? (T.Q0 ? T.Q0[ri * N + c] : (ri == c ? 1.0 : 0.0)) : 0.0;
  }
  auto row8_sum = [](double v) {
    v = dpp_add<0xb1, 0xf>(v);
    v = dpp_add<0x4e, 0xf>(v);
    return dpp_add<0x141, 0xf>(v);
  };
  double* col = lo;  // column k of the current A (64 rows)
  double* pb = wv;   // p (64 rows)
  if (red && g == 0) col[ri] = ar[0];
  if (tid == 0) *fail = 0;
  __syncthreads();
  for (int k = 0; k + 2 < N; ++k) {
    double uj[8], ui = 0.0, tau = 0.0, p = 0.0;
    if (red) {  // reflector of column k from col
      double s2[2] = {0.0, 0.0};  // two partial sums: shorter dependency chains
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int c = g + 8 * u;
        uj[u] = (c > k && c < N) ? col[c] : 0.0;
        s2[u & 1] = fma(uj[u], uj[u], s2[u & 1]);
      }
      const double ss = row8_sum(s2[0] + s2[1]);
      const double ui0 = (ri > k && ri < N) ? col[ri] : 0.0;
      const double x0 = col[k + 1];
      double alpha = 0.0;
      if (ss > 0.0) {
        const double nx = sqrt(ss);
        alpha = x0 >= 0.0 ? -nx : nx;
        const double v0 = x0 - alpha;
        tau = recip_nr(nx * (nx + fabs(x0)));
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (g + 8 * u == k + 1) uj[u] = v0;
        ui = ri == k + 1 ? v0 : ui0;
      }
      if (tid == 0) e[k] = alpha;
      // p = tau A u and Q <- Q - tau (Q u) u^T
      double sa[2] = {0.0, 0.0}, sq[2] = {0.0, 0.0};
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        sa[u & 1] = fma(ar[u], uj[u], sa[u & 1]);
        sq[u & 1] = fma(qr[u], uj[u], sq[u & 1]);
      }
      p = tau * row8_sum(sa[0] + sa[1]);
      const double y = tau * row8_sum(sq[0] + sq[1]);
#pragma unroll
      for (int u = 0; u < 8; ++u) qr[u] = fma(-y, uj[u], qr[u]);
      if (g == 0) pb[ri] = p;
    }
    __syncthreads();
    if (red) {  // K = tau (u^T p) / 2, w = p - K u, A <- A - u w^T - w u^T
      double pj[8], kv[2] = {0.0, 0.0};
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        pj[u] = pb[g + 8 * u];
        kv[u & 1] = fma(uj[u], pj[u], kv[u & 1]);
      }
      const double K = 0.5 * tau * row8_sum(kv[0] + kv[1]);
      const double wi = fma(-K, ui, p);
#pragma unroll
      for (int u = 0; u < 8; ++u) ar[u] -= fma(ui, fma(-K, uj[u], pj[u]), wi * uj[u]);
#pragma unroll
      for (int u = 0; u < 8; ++u)  // constant indices keep ar in registers
        if (g + 8 * u == k + 1) col[ri] = ar[u];
    }
    __syncthreads();
  }
  // diagonal (final once its step is past), the last off-diagonal, and Q into
  // A (the input is no longer needed); Z is free for the eigenvectors of the
  // tridiagonal matrix
  if (red) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int c = g + 8 * u;
      if (ri < N && c == ri) d[ri] = ar[u];
      if (ri == N - 1 && c == N - 2) e[N - 2] = ar[u];
      if (ri < N && c < N) A[ri * ld + c] = qr[u];
    }
  }
  __syncthreads();
  // 2. splits and the block of every position
  if (wave == 0) {
    const bool sp = lane < N - 1 && fabs(e[lane]) <= 1e-12 * (fabs(d[lane]) + fabs(d[lane + 1]));
    if (sp) e[lane] = 0.0;
    const unsigned long long mask = __ballot(sp);  // bit i: split after position i
    if (lane == 0) fail[1] = mask == 0ull ? 1 : 0;
    if (lane < N) {
      const unsigned long long below = lane > 0 ? (mask & ((1ull << lane) - 1)) : 0ull;
      blo[lane] = below ? 64 - __clzll(below) : 0;   // highest split below lane, + 1
      const unsigned long long at = mask >> lane;
      bhi[lane] = at ? lane + __ffsll(at) : N;       // lowest split at or above lane, + 1
    }
    wave_sync();
    // Gershgorin interval of slot lane's block (the whole matrix's when it
    // does not split: one wavefront reduction)
    const double rad = lane < N ? (lane > 0 ? fabs(e[lane - 1]) : 0.0) + (lane + 1 < N ? fabs(e[lane]) : 0.0) : 0.0;
    const double gl = -wave_max(lane < N ? rad - d[lane] : -1e300);
    const double gh = wave_max(lane < N ? d[lane] + rad : -1e300);
    if (lane < N) {
      double a = gl, b = gh;
      if (mask != 0ull) {
        a = 1e300;
        b = -1e300;
        for (int i = blo[lane]; i < bhi[lane]; ++i) {
          const double r = (i > blo[lane] ? fabs(e[i - 1]) : 0.0) + (i + 1 < bhi[lane] ? fabs(e[i]) : 0.0);
          a = fmin(a, d[i] - r);
          b = fmax(b, d[i] + r);
        }
      }
      const double pad = 4.0 * 2.220446049250313e-16 * fmax(fabs(a), fabs(b)) + 1e-300;
      lo[lane] = a - pad;
      hi[lane] = b + pad;
    } else {
      lo[lane] = 0.0;
      hi[lane] = 0.0;
      blo[lane] = 0;
      bhi[lane] = 0;
      d[lane] = 0.0;
    }
    wv[lane] = (lane > 0 && lane < N) ? e[lane - 1] * e[lane - 1] : 0.0;  // e_(i-1)^2, 0 at splits
  }
  __syncthreads();
  // 3. multisection: wave q < TRI_SECT evaluates section point q of every
  // slot (lane), TRI_ROUNDS rounds narrow each interval (TRI_SECT + 1)-fold
  {
    const int b0 = blo[lane], b1 = bhi[lane];
    const bool whole = fail[1] != 0;  // no splits
    for (int round = 0; round < TRI_ROUNDS; ++round) {
      if (wave < TRI_SECT) {
        const double x = lo[lane] + (wave + 1) * ((hi[lane] - lo[lane]) * (1.0 / (TRI_SECT + 1)));
        cnt[wave * 64 + lane] = whole ? sturm_count_whole(d, wv, N, x) : sturm_count_all(d, wv, N, b0, b1, x);
      }
      __syncthreads();
      if (wave == 0 && lane < N) {
        const int j = lane - b0;
        const double a = lo[lane], w = (hi[lane] - a) * (1.0 / (TRI_SECT + 1));
        // the first section point whose count exceeds j (branch-free)
        int first = TRI_SECT;
#pragma unroll
        for (int q = TRI_SECT - 1; q >= 0; --q) first = cnt[q * 64 + lane] > j ? q : first;
        const double na = first > 0 ? a + first * w : a;
        const double nb = first < TRI_SECT ? a + (first + 1) * w : hi[lane];
        lo[lane] = na;
        hi[lane] = nb;
      }
      __syncthreads();
    }
  }
  // 4. ascending order of the slots, then the eigenvectors
  if (wave == 0) {  // slot q's eigenvalue from lane q by readlane
    const double lk = lane < N ? 0.5 * (lo[lane] + hi[lane]) : 0.0;
    int r = 0;
    for (int q = 0; q < N; ++q) {
      const double lq = readlane_f64(lk, q);
      r += (lq < lk) || (lq == lk && q < lane);
    }
    if (lane < N) {
      T.rank[lane] = r;
      lam[lane] = lk;
    }
  }
  __syncthreads();
  if (wave == 0 && lane < N) T.eig[T.rank[lane]] = lam[lane];
  for (int x = tid; x < N * ld; x += blockDim.x) Z[x] = 0.0;
  __syncthreads();
  if (T.W2 != nullptr) {
    // all slots in one pass (lane = slot): 1 / Dp in W, 1 / Dm in W2 at the
    // odd stride ld.  Without splits every slot's block is [0, N) and the
    // loops run unpredicated.
    const bool whole = fail[1] != 0;  // no splits
    if (wave == 0) {
      const bool act = lane < N;
      const int a0 = act ? blo[lane] : 0, a1 = act ? bhi[lane] : 0;
      const double lk = act ? lam[lane] : 0.0;
      const int col = act ? T.rank[lane] : 0;
      double* rdp = W + (act ? lane : 0) * ld;
      double* rdm = T.W2 + (act ? lane : 0) * ld;
      if (whole) {
        const double sc = wave_max(act ? fabs(d[lane]) + (lane < N - 1 ? fabs(e[lane]) : 0.0) : 0.0);
        twisted_whole(d, e, N, ld, 2.220446049250313e-16 * sc + 1e-300, lk, act, col, rdp, rdm, Z, dot,
                      kpart);
      } else {
        twisted_slot<true>(d, e, N, ld, a0, a1, lk, act, col, rdp, rdm, Z);
      }
    }
    __syncthreads();
    if (whole) {  // normalise the columns
      for (int x = tid; x < N * N; x += blockDim.x) {
        const int i = x / N, c = x - i * N;
        Z[i * ld + c] *= dot[c];
      }
      __syncthreads();
    }
  }
  // slots per pass, each with N doubles of 1 / Dp and of 1 / Dm at an odd
  // stride (distinct LDS banks across the lanes)
  const int stride = 2 * N + 1;
  const int bs = T.W2 != nullptr ? N : (T.wcap / stride < 32 ? T.wcap / stride : 32);
  for (int b0 = T.W2 != nullptr ? N : 0; b0 < N; b0 += bs) {
    const int k = b0 + lane;
    if (wave == 0 && lane < bs && k < N) {
      const int a0 = blo[k], a1 = bhi[k];
      const double lk = lam[k];
      double* rdp = W + lane * stride;  // 1 / Dp (rows a0..a1 at offsets 0..)
      double* rdm = rdp + N;            // 1 / Dm
      double sc = 0.0;
      for (int i = a0; i < a1; ++i) sc = fmax(sc, fabs(d[i]) + (i < a1 - 1 ? fabs(e[i]) : 0.0));
      const double tiny = 2.220446049250313e-16 * sc + 1e-300;
      auto guard = [&](double v) { return fabs(v) >= tiny ? v : (v < 0.0 ? -tiny : tiny); };
      double dp = guard(d[a0] - lk);
      double r_prev = recip_nr(dp);
      rdp[0] = r_prev;
      for (int i = a0 + 1; i < a1; ++i) {
        const double ei = e[i - 1];
        dp = guard((d[i] - lk) - ei * ei * r_prev);
        r_prev = recip_nr(dp);
        rdp[i - a0] = r_prev;
      }
      // bottom-up, with gamma_i = Dm_i - e_(i-1)^2 / Dp_(i-1)
      double dm = guard(d[a1 - 1] - lk);
      double rnext = recip_nr(dm);
      rdm[a1 - 1 - a0] = rnext;
      int r = a1 - 1;
      double gbest = fabs(a1 - 1 > a0 ? dm - e[a1 - 2] * e[a1 - 2] * rdp[a1 - 2 - a0] : dm);
      for (int i = a1 - 2; i >= a0; --i) {
        const double ei = e[i];
        dm = guard((d[i] - lk) - ei * ei * rnext);
        rnext = recip_nr(dm);
        rdm[i - a0] = rnext;
        const double g = fabs(i > a0 ? dm - e[i - 1] * e[i - 1] * rdp[i - 1 - a0] : dm);
        if (g < gbest) {
          gbest = g;
          r = i;
        }
      }
      const int col = T.rank[k];
      double z = 1.0, nrm = 1.0;
      Z[r * ld + col] = 1.0;
      for (int i = r - 1; i >= a0; --i) {
        z = -e[i] * rdp[i - a0] * z;
        Z[i * ld + col] = z;
        nrm += z * z;
      }
      z = 1.0;
      for (int i = r + 1; i < a1; ++i) {
        z = -e[i - 1] * rdm[i - a0] * z;
        Z[i * ld + col] = z;
        nrm += z * z;
      }
      const double inv = 1.0 / sqrt(nrm);
      for (int i = a0; i < a1; ++i) Z[i * ld + col] *= inv;
    }
    __syncthreads();
  }
  // 5. clusters: modified Gram-Schmidt in ascending order (one wave, lane = row)
  if (wave == 0) {
    const double sc = wave_max(lane < N ? fabs(d[lane]) + (lane < N - 1 ? fabs(e[lane]) : 0.0) : 0.0);
    const double tau = 1e-6 * sc;
    // columns closer than tau to their predecessor (none: nothing to do)
    const bool any = __ballot(lane >= 1 && lane < N && !(T.eig[lane] - T.eig[lane - 1] >= tau)) != 0ull;
    int c0 = 0;
    for (int c = 1; any && c < N; ++c) {
      if (T.eig[c] - T.eig[c - 1] >= tau) {
        c0 = c;
        continue;
      }
      double zc = lane < N ? Z[lane * ld + c] : 0.0;
      for (int q = c0; q < c; ++q) {
        const double zq = lane < N ? Z[lane * ld + q] : 0.0;
        zc -= wave_sum(zq * zc) * zq;
      }
      const double nrm = sqrt(wave_sum(zc * zc));
      if (!(nrm >= 0.5)) {
        if (lane == 0) *fail = 1;
        break;
      }
      if (lane < N) Z[lane * ld + c] = zc / nrm;
      wave_sync();
    }
  }
  __syncthreads();
  if (*fail) return false;
  // 6. eigenvectors Q Z (Q in A), by way of W
  lds_matmul<false, false>(A, Z, W, N, ld, ld);
  __syncthreads();
  for (int x = tid; x < N * ld; x += blockDim.x) Z[x] = W[x];
  __syncthreads();
  // finite results only
  bool bad = false;
  for (int x = tid; x < N * N; x += blockDim.x) {
    const int i = x / N;
    bad |= !isfinite(Z[i * ld + x - i * N]);
  }
  if (tid < N) bad |= !isfinite(T.eig[tid]);
  if (bad) *fail = 1;
  __syncthreads();
  return *fail == 0;
}

// From the symmetric matrix (L.A) and the initial rotation (L.U, identity or
// a warm start): Jacobi, ascending ranks of the eigenvalues, the eigenvectors
// in the AO basis left in L.A with columns unsorted -- C = L^-T U (Cb = L^-1,
// in the orthonormal basis) or C = C_prev U (cbasis: Cb = C_prev, in the
// previous step's eigenbasis) -- and C_occ C_occ^T (= D / 2) left in L.U
// (leading dimension ld).  Wout (the next step's warm start) receives U
// (store_c false) or C (store_c true).
__device__ __forceinline__ int eig_density(const ScfLds& L, int N, int nocc, const double* __restrict__ Cb,
                                  double* __restrict__ Wout, bool solved = false, bool store_c = false,
                                  bool cbasis = false) {
  const int ld = lds_ld(N), tid = threadIdx.x;
  double *A = L.A, *U = L.U, *B = L.B;
  const int sweeps = solved ? 0 : parallel_jacobi(A, U, N, ld, L.rp, L.rq, L.rc, L.rs, L.red);
  if (Wout && !store_c)  // eigenvectors in the orthonormal basis
    for (int x = tid; x < N * N; x += blockDim.x) {
      const int i = x / N;
      Wout[x] = U[i * ld + x - i * N];
    }
  if (!solved)
    for (int j = tid; j < N; j += blockDim.x) L.eig[j] = A[j * ld + j];
  lds_load(B, Cb, N, ld);
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
  if (cbasis)
    lds_matmul<false, false>(B, U, A, N, ld, ld);  // A = C_prev U (columns unsorted)
  else
    lds_matmul<true, false>(B, U, A, N, ld, ld);  // A = L^-T U (columns unsorted)
  __syncthreads();
  if (Wout && store_c)  // eigenvectors in the AO basis
    for (int x = tid; x < N * N; x += blockDim.x) {
      const int i = x / N;
      Wout[x] = A[i * ld + x - i * N];
    }
  // occupied columns into B, then D = 2 B A^T into B via U
  for (int x = tid; x < N * N; x += blockDim.x) {
    const int i = x / N, j = x - i * N;
    B[i * ld + j] = L.rank[j] < nocc ? A[i * ld + j] : 0.0;
  }
  __syncthreads();
  lds_matmul_sym<false, true>(B, A, U, N, ld, ld);  // U = C_occ C^T
  __syncthreads();
  return sweeps;
}

// First row of B^-1 for the leading nv x nv block (nv <= 10) of the row-major
// 10 x 10 matrix Bm, by Gauss-Jordan elimination with partial pivoting in one
// wavefront: lane c < 2 nv holds column c of [B | I] in registers, pivots and
// multipliers travel by readlane.  Writes c[j] = (B^-1)[0, j] for j < nv and
// 0 for nv <= j < 10, and returns true, when kappa = |B|_F |B^-1|_F is finite
// and below 4.5e11: then every eigenvalue w of B has |w| >= 1 / |B^-1|_F >
// 1e2 (100 eps |B|_F) >= 1e2 x the pseudo-inverse cutoff 100 eps max |w|, so
// B^-1 is the pseudo-inverse.  Otherwise returns false with c untouched.
// Called by a whole wavefront (its return value is wavefront-uniform).
__device__ __forceinline__ bool diis_inverse_row(const double* __restrict__ Bm, int nv, double* __restrict__ c) {
  const int lane = threadIdx.x & 63;
  double a[10];
  double nb = 0.0;
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    double v = 0.0;
    if (i < nv) {
      if (lane < nv)
        v = Bm[i * 10 + lane];
      else if (lane < 2 * nv)
        v = (lane - nv == i) ? 1.0 : 0.0;
    }
    a[i] = v;
    if (lane < nv) nb += v * v;
  }
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    if (k < nv) {  // wavefront-uniform
      int p = k;
      double best = -1.0;
#pragma unroll
      for (int i = k; i < 10; ++i) {
        const double m = fabs(a[i]);
        if (i < nv && m > best) {
          best = m;
          p = i;
        }
      }
      p = __builtin_amdgcn_readlane(p, k);
#pragma unroll
      for (int i = k + 1; i < 10; ++i)
        if (i == p) {
          const double t = a[i];
          a[i] = a[k];
          a[k] = t;
        }
      a[k] *= 1.0 / readlane_f64(a[k], k);
#pragma unroll
      for (int i = 0; i < 10; ++i)
        if (i != k && i < nv) a[i] = fma(-readlane_f64(a[i], k), a[k], a[i]);
    }
  }
  double ni = 0.0;
  if (lane >= nv && lane < 2 * nv)
#pragma unroll
    for (int i = 0; i < 10; ++i) ni += a[i] * a[i];
  const double k2 = wave_sum(nb) * wave_sum(ni);
  if (!(k2 < 4.5e11 * 4.5e11)) return false;
  if (lane >= nv && lane < 2 * nv) c[lane - nv] = a[0];
  if (lane >= 20 && lane < 30 && lane - 20 >= nv) c[lane - 20] = 0.0;
  return true;
}

// Sums of a, b and e and maxima of c and d over the workgroup (at most 16
// wavefronts; red holds 80 doubles); every thread receives all five.
__device__ inline void block_reduce5(double& a, double& b, double& c, double& d, double& e, double* red) {
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_max(c);
  d = wave_max(d);
  e = wave_sum(e);
  __syncthreads();
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w] = a;
    red[16 + w] = b;
    red[32 + w] = c;
    red[48 + w] = d;
    red[64 + w] = e;
  }
  __syncthreads();
  a = combine_sum(red, nw);
  b = combine_sum(red + 16, nw);
  c = combine_max(red + 32, nw);
  d = combine_max(red + 48, nw);
  e = combine_sum(red + 64, nw);
  __syncthreads();
}

constexpr int OA_MAX = 10;           // refinement updates before giving up
constexpr double OA_ENTRY = 3e-6;    // off-diagonal share of |A'|^2 up to which the refinement runs
constexpr double OA_TAU = 0.2;       // |first-order correction| above which a pair is near-degenerate

// Refinement of an approximate eigenbasis X (L.U, orthonormal to rounding) of
// the symmetric A (L.M), after Ogita and Aishima, "Iterative refinement for
// symmetric eigenvalue decomposition" (Japan J. Indust. Appl. Math. 35, 2018).
// Each iteration, on FP64 MFMA products:
//   G = X^T X (R = I - G),  S = X^T A X,  lam_i = S_ii / G_ii,
//   E_ij = (S_ij + lam_j R_ij) / (lam_j - lam_i),   E_ii = R_ii / 2,
//   X <- X + X E.
// Convergence is quadratic while every |E_ij| is small.  A pair whose
// first-order correction exceeds OA_TAU is near-degenerate: when it is the
// only such pair of both its rows, it takes the exact 2 x 2 Jacobi rotation
// instead, E_ij = -E_ji = tan(theta) with tan(2 theta) = 2 S_ij / (lam_j -
// lam_i) (bounded by 1), plus R_ij / 2; inside a larger group, and when the
// two eigenvalues agree to 1e-13, it is only re-orthogonalised, E_ij = R_ij /
// 2 (its invariant subspace still converges, and for an exactly degenerate
// group any basis of it is an eigenbasis).  Converged when the off-diagonal share of
// |S|^2 is at most 1e-18 (the Jacobi criterion) and |R| <= 1e-13, or after an
// update with every |E_ij| <= 1e-9 (the next S and R would be second order in
// E): returns the number of updates, with X in L.U and the eigenvalues in
// L.eig.  Returns -1 (X then undefined) after OA_MAX updates or on a
// non-finite update.  s_ready: L.A already holds S = X^T A X.  L.A, L.B and
// L.rp are scratch.  x_id: X is the identity to start, so the first X^T X is
// the identity without a product.  N <= 64, 1024 threads.
__device__ __forceinline__ int oa_refine(const ScfLds& L, int N, bool s_ready, bool x_id = false) {
  const int ld = lds_ld(N), tid = threadIdx.x;
  double *M = L.M, *X = L.U, *S = L.A, *G = L.B;
  int* nbig = L.rp;  // near-degenerate pairs per row
  if (tid < 64) nbig[tid] = 0;
  for (int k = 0;; ++k) {
    if (k > 0 || !s_ready) {
      lds_matmul<false, false>(M, X, G, N, ld, ld);  // G = A X
      __syncthreads();
      lds_matmul_sym<true, false>(X, G, S, N, ld, ld);  // S = X^T A X
      __syncthreads();
    }
    if (k == 0 && x_id)
      for (int x = tid; x < N * ld; x += blockDim.x) G[x] = (x / ld == x % ld) ? 1.0 : 0.0;
    else
      lds_matmul_sym<true, false>(X, X, G, N, ld, ld);  // G = X^T X
    __syncthreads();
    if (tid < N) L.eig[tid] = S[tid * ld + tid] / G[tid * ld + tid];
    __syncthreads();
    // one pass over the pairs i <= j: convergence measures and the first-order
    // corrections E (in place of S); near-degenerate pairs are only counted
    // here (their S entries stay for the second pass)
    double off = 0.0, tot = 0.0, flags = 0.0, rmax = 0.0, emax = 0.0;
    // pairs i <= j folded into lines of N + 1: line k holds row k from the
    // diagonal (N - k pairs) then row N - 1 - k from the diagonal (k + 1
    // pairs); for odd N the middle row is a last, shorter line
    const float inv_line = 1.0f / (float)(N + 1);
    for (int x = tid; x < N * (N + 1) / 2; x += blockDim.x) {
      const int k = (int)(((float)x + 0.5f) * inv_line), c = x - k * (N + 1);
      const int i = c < N - k ? k : N - 1 - k, j = c < N - k ? k + c : c - 1;
      const double sij = S[i * ld + j];
      if (i == j) {
        const double rii = 1.0 - G[i * ld + i];
        tot += sij * sij;
        rmax = fmax(rmax, fabs(rii));
        S[i * ld + i] = 0.5 * rii;
        emax = fmax(emax, fabs(0.5 * rii));
        continue;
      }
      const double sji = S[j * ld + i];
      const double rij = -G[i * ld + j], rji = -G[j * ld + i];
      off += sij * sij + sji * sji;
      tot += sij * sij + sji * sji;
      rmax = fmax(rmax, fmax(fabs(rij), fabs(rji)));
      const double sym = 0.5 * (sij + sji);
      const double li = L.eig[i], lj = L.eig[j], dl = lj - li;
      const double nij = sym + lj * rij, nji = sym + li * rji;
      if (fabs(dl) <= 1e-13 * fmax(fabs(li), fabs(lj)) || fabs(nij) > OA_TAU * fabs(dl)) {
        atomicAdd(&nbig[i], 1);
        atomicAdd(&nbig[j], 1);
        flags += 1.0;
        continue;
      }
      const double eij = nij / dl, eji = -nji / dl;
      S[i * ld + j] = eij;
      S[j * ld + i] = eji;
      flags += (fabs(eij) <= 1.5 && fabs(eji) <= 1.5) ? 0.0 : 1e6;  // non-finite entries
      emax = fmax(emax, fmax(fabs(eij), fabs(eji)));
    }
    block_reduce5(off, tot, rmax, emax, flags, L.red);  // also publishes nbig and E
    if (off <= 1e-18 * tot && rmax <= 1e-13) return k;
    if (k == OA_MAX || !(flags < 1e6)) return -1;
    if (flags > 0.0) {  // near-degenerate pairs
      for (int x = tid; x < N * N; x += blockDim.x) {
        const int i = x / N, j = x - i * N;
        if (i >= j) continue;
        const double sym = 0.5 * (S[i * ld + j] + S[j * ld + i]);
        const double rij = -G[i * ld + j], rji = -G[j * ld + i];
        const double li = L.eig[i], lj = L.eig[j], dl = lj - li;
        const bool degen = fabs(dl) <= 1e-13 * fmax(fabs(li), fabs(lj));
        if (!degen && !(fabs(sym + lj * rij) > OA_TAU * fabs(dl))) continue;
        double eij = 0.5 * rij, eji = 0.5 * rji;
        if (!degen && nbig[i] == 1 && nbig[j] == 1) {  // isolated pair: exact rotation (+ re-orthogonalisation)
          const double h = 0.5 * dl;
          const double t = copysign(1.0, h) * sym / (fabs(h) + sqrt(h * h + sym * sym));
          eij += t;
          eji -= t;
        }
        S[i * ld + j] = eij;
        S[j * ld + i] = eji;
      }
      __syncthreads();
      if (tid < 64) nbig[tid] = 0;  // read again only after the next products
    }
    // an update of at most 1e-9 leaves the off-diagonal share and X^T X - I at
    // the rounding level (both are second order in E): no further check
    const bool last = flags == 0.0 && emax <= 1e-9;
    const double* XE = S;  // X = I (first update of x_id): X E = E exactly
    if (k > 0 || !x_id) {
      lds_matmul<false, false>(X, S, G, N, ld, ld);  // G = X E
      __syncthreads();
      XE = G;
    }
    for (int x = tid; x < N * N; x += blockDim.x) {
      const int i = x / N, j = x - i * N;
      X[i * ld + j] += XE[i * ld + j];
    }
    __syncthreads();
    if (last) return k + 1;
  }
}

// A = L^-1 H L^-T, exactly symmetric, from H in A and L^-1 in B (LDS); U is
// scratch.  (Rounding of the FP32 Fock terms and of the transform leaves
// O(eps) asymmetry that Jacobi rotations cannot remove.)
__device__ inline void orthonormal_basis(const ScfLds& L, int N) {
  const int ld = lds_ld(N), tid = threadIdx.x;
  double *A = L.A, *U = L.U, *B = L.B;
  lds_matmul<false, true>(A, B, U, N, ld, ld);  // U = H L^T
  __syncthreads();
  lds_matmul_sym<false, false>(B, U, A, N, ld, ld);  // A = L U
  __syncthreads();
  for (int x = tid; x < N * N; x += blockDim.x) {
    const int i = x / N, j = x - i * N;
    if (i < j) {
      const double v = 0.5 * (A[i * ld + j] + A[j * ld + i]);
      A[i * ld + j] = v;
      A[j * ld + i] = v;
    }
  }
  __syncthreads();
}

// One SCF step for one molecule by the whole workgroup (see pscf_scf_step).
// Optional Dt receives D converted to T (the precision of the next J/K / XC),
// optional dDt the change of D in T (for an incremental J/K build).
template <typename T, typename TJ>
__device__ __forceinline__ void scf_step_body(const ScfLds& L, int it, int N, int nocc, const double* __restrict__ Hcore,
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
  // the previous step's eigenvectors: C = L^-T U in the AO basis where the
  // eigen-decomposition works in their basis (N <= 64 with 1024 threads), else
  // U in the orthonormal basis
  double* Wprev = Bst + 100;
  double* flags = Wprev + NN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = blockDim.x >> 6;
  const bool warm = flags[0] != 0.0;
  const bool small = L.M != nullptr && N <= 64 && blockDim.x == 1024;
  const bool cb = small && warm;  // eigenbasis of the previous step

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
    auto masked_B = [&]() {
      for (int x = lane; x < 100; x += 64) {
        const int i = x / 10, j = x % 10;
        const bool mi = i == 0 || i - 1 < nvalid, mj = j == 0 || j - 1 < nvalid;
        L.Bm[x] = (mi && mj) ? L.Bst[x] : 0.0;
      }
    };
    if (wave == 0) masked_B();
    __syncthreads();
    // c = pinv(B)[0]: the inverse of the valid block while B is far from
    // singular (every step of a converging SCF run), else from the
    // eigen-decomposition of the 10 x 10 system
    if (wave == 0) {
      const bool inv = diis_inverse_row(L.Bm, nvalid + 1, L.cvec);
      if (lane == 0) L.cvec[15] = inv ? 1.0 : 0.0;
    }
    __syncthreads();
    const bool inv_ok = L.cvec[15] != 0.0;
    // the tridiagonal solver (all threads; eigenvalues into red, ranks into
    // rp, U as its scratch), else one-wavefront Jacobi
    const bool tri_ok = !inv_ok && L.tri != nullptr && blockDim.x == 1024 && N * ld >= 20 &&
                        tri_eigh(TriBufs{L.Bm, L.Bv, L.U, L.red, L.rp, L.tri, 10, N * ld, nullptr, nullptr}, 10);
    if (wave == 0 && !inv_ok) {
      if (!tri_ok) {
        masked_B();
        wave_sync();
        wave_jacobi(L.Bm, L.Bv, 10, L.rp, L.rq, L.rc, L.rs);
      }
      double wmax = 0.0;
      for (int i = 0; i < 10; ++i) wmax = fmax(wmax, fabs(tri_ok ? L.red[i] : L.Bm[i * 10 + i]));
      const double cut = 10.0 * 10.0 * 2.220446049250313e-16 * wmax;
      // pinv(B)[0, j] = sum_k V[0,k] V[j,k] / w_k over |w_k| > cut
      if (lane < 10) {
        double acc = 0.0;
        for (int k = 0; k < 10; ++k) {
          const double wk = tri_ok ? L.red[k] : L.Bm[k * 10 + k];
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
  lds_load(B, cb ? Wprev : Linv, N, ld);
  __syncthreads();
  for (int x = tid; x < NN; x += blockDim.x) {
    const int i = x / N;
    Hout[x] = A[i * ld + x - i * N];
  }
  if (!cb) orthonormal_basis(L, N);
  // Eigen-decomposition (N <= 64 with 1024 threads):
  //  * first step (no previous eigenvectors): the tridiagonal solver on A =
  //    L^-1 H L^-T;
  //  * later steps: in the previous eigenbasis, A' = C_prev^T H C_prev (= U_prev^T
  //    A U_prev: two products from H instead of four).  Where its off-diagonal
  //    share of |A'|^2 is at most OA_ENTRY (the SCF's later steps), the
  //    refinement from X = I (a few MFMA products per iteration); otherwise
  //    the tridiagonal solver on A'.  C = C_prev X.
  // Cyclic Jacobi (from the identity, or from U_prev for N > 64) takes over
  // when either fails, and serves N > 64.
  auto start = [&]() {  // N > 64: A = Uprev^T A Uprev, U = Uprev
    lds_load(B, Wprev, N, ld);
    __syncthreads();
    lds_matmul<false, false>(A, B, U, N, ld, ld);  // U = A Uprev
    __syncthreads();
    lds_matmul_sym<true, false>(B, U, A, N, ld, ld);  // A = Uprev^T U
    __syncthreads();
    for (int x = tid; x < N * ld; x += blockDim.x) U[x] = B[x];
    __syncthreads();
  };
  auto identity_u = [&]() {
    for (int x = tid; x < N * ld; x += blockDim.x) U[x] = (x / ld == x % ld) ? 1.0 : 0.0;
    __syncthreads();
  };
  auto start_c = [&]() {  // H in A, C_prev in B: A = C_prev^T H C_prev, U = I
    lds_matmul<false, false>(A, B, U, N, ld, ld);  // U = H C_prev
    __syncthreads();
    lds_matmul_sym<true, false>(B, U, A, N, ld, ld);  // A = C_prev^T U
    __syncthreads();
    identity_u();
  };
  auto rebuild = [&]() {  // A from H after the tridiagonal solver overwrote it
    lds_load(A, Hout, N, ld);
    lds_load(B, Linv, N, ld);
    __syncthreads();
    orthonormal_basis(L, N);
  };
  int path = 0;  // 1 tridiagonal, 2 refinement
  int refined = -1;
  if (!warm) {
    if (small) {
      if (tri_eigh(TriBufs{L.A, L.U, L.B, L.eig, L.rank, L.tri, ld, N * ld, L.M, nullptr}, N))
        path = 1;
      else
        rebuild();
    }
    if (path == 0) identity_u();
  } else if (small) {
    start_c();
    for (int x = tid; x < N * ld; x += blockDim.x) L.M[x] = A[x];  // A', kept
    __syncthreads();
    double off = 0.0, tot = 0.0, z0 = 0.0, z1 = 0.0, z2 = 0.0;
    for (int x = tid; x < NN; x += blockDim.x) {
      const int i = x / N, j = x - i * N;
      const double v = A[i * ld + j] * A[i * ld + j];
      tot += v;
      if (i != j) off += v;
    }
    block_reduce5(off, tot, z0, z1, z2, L.red);
    if (off <= OA_ENTRY * tot) {
      refined = oa_refine(L, N, true, true);  // X = I: S of the first iteration is A' itself
      if (refined >= 0) {
        path = 2;
      } else {  // Jacobi on A' from the identity
        for (int x = tid; x < N * ld; x += blockDim.x) A[x] = L.M[x];
        __syncthreads();
        identity_u();
      }
    } else if (tri_eigh(TriBufs{L.A, L.U, L.B, L.eig, L.rank, L.tri, ld, N * ld, L.M, nullptr}, N)) {
      path = 1;  // U = eigenvectors of A'
    } else {  // A' again (the solver overwrote A, B and M)
      lds_load(A, Hout, N, ld);
      lds_load(B, Wprev, N, ld);
      __syncthreads();
      start_c();
    }
  } else {
    start();
  }
  int sweeps = 0;
  if (path == 0) {  // one call site: the sweep loop is compiled inline once
    sweeps = parallel_jacobi(A, U, N, ld, L.rp, L.rq, L.rc, L.rs, L.red);
    for (int j = tid; j < N; j += blockDim.x) L.eig[j] = A[j * ld + j];
  }
  __syncthreads();
  eig_density(L, N, nocc, cb ? Wprev : Linv, Wprev, true, small, cb);
  if (tid == 0) {
    flags[0] = 1.0;
    // diagnostic: -1 tridiagonal solver, 100 (k + 1) refinement of k updates,
    // s >= 0 Jacobi sweeps
    flags[1] = path == 1 ? -1.0 : (path == 2 ? 100.0 * (refined + 1) : (double)sweeps);
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
  L.tri = nullptr;
  L.M = nullptr;
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
