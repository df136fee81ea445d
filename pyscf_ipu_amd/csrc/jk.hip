// Sparse symmetric J/K contraction on gfx950.
//
// Reference: sparse_symmetric_einsum (pyscf_ipu/nanoDFT/sparse_symmetric_ERI.py:86-137)
// runs 16 serial gather(dm) -> multiply -> scatter-add passes over the unique
// ERIs, with flat indices from the 32-case switch of compute_indices.cpp:54-95.
// Here one pass does all 16 patterns per ERI with dm and the accumulator
// resident in LDS (ds_add atomics), each workgroup writing a private partial
// that a second kernel reduces in a fixed order:
//   * jk_sparse: the drop-in (explicit int16 indices, multiplicity-divided
//     values, any dm); the 8 J patterns fold into 4 gathers / 4 scatters;
//   * jk_s8: the packed s8 array of pscf_eri_s8, indices regenerated from the
//     element position (4 B of HBM traffic per ERI), dm symmetric, so J and K
//     fold into one accumulator M with dJK = M + M^T (6 scatters per ERI).
#include <mutex>

#include "batch.h"
#include "common.h"
#include "reduce.h"
#include "wave.h"

namespace pscf {

static int max_lds_bytes() { return device_max_lds(); }

constexpr int JK_THREADS = 256;
constexpr int JK_MAX_BLOCKS = 240;

// ------------------------------------------------------------- sparse form ---
template <typename T, bool LDS>
__global__ void __launch_bounds__(JK_THREADS) jk_sparse_kernel(const T* __restrict__ eri,
                                                               const int16_t* __restrict__ idx, int64_t nnz,
                                                               const T* __restrict__ dm_g, int N, T kscale,
                                                               T* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int NN = N * N;
  const T* dm = dm_g;
  T* acc = part;  // global-atomic mode accumulates into out directly
  if constexpr (LDS) {
    T* sdm = reinterpret_cast<T*>(smem);
    acc = sdm + NN;
    for (int x = threadIdx.x; x < NN; x += blockDim.x) {
      sdm[x] = dm_g[x];
      acc[x] = T(0);
    }
    __syncthreads();
    dm = sdm;
  }
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += stride) {
    const uint2 packed = reinterpret_cast<const uint2*>(idx)[e];
    const int i = (int)(packed.x & 0xffffu), j = (int)(packed.x >> 16);
    const int k = (int)(packed.y & 0xffffu), l = (int)(packed.y >> 16);
    // out-of-range indices are dropped, like the reference's FILL_OR_DROP gather/scatter
    if (i >= N || j >= N || k >= N || l >= N) continue;
    const T v = eri[e];
    const int ij = i * N + j, ji = j * N + i, kl = k * N + l, lk = l * N + k;
    // J: patterns 0-7 gather {ij, ji, kl, lk}, scatter {kl, lk, ij, ji}
    const T tj1 = v * (dm[ij] + dm[ji]);
    const T tj2 = v * (dm[kl] + dm[lk]);
    atomicAdd(acc + kl, tj1);
    atomicAdd(acc + lk, tj1);
    atomicAdd(acc + ij, tj2);
    atomicAdd(acc + ji, tj2);
    // K: patterns 16-23 gather / 24-31 scatter, scaled by -hyb/2
    const T vk = v * kscale;
    const int il = i * N + l, jl = j * N + l, ik = i * N + k, jk = j * N + k;
    const int kj = k * N + j, lj = l * N + j, ki = k * N + i, li = l * N + i;
    atomicAdd(acc + il, vk * dm[kj]);
    atomicAdd(acc + jl, vk * dm[ki]);
    atomicAdd(acc + ik, vk * dm[lj]);
    atomicAdd(acc + jk, vk * dm[li]);
    atomicAdd(acc + kj, vk * dm[il]);
    atomicAdd(acc + lj, vk * dm[ik]);
    atomicAdd(acc + ki, vk * dm[jl]);
    atomicAdd(acc + li, vk * dm[jk]);
  }
  if constexpr (LDS) {
    __syncthreads();
    T* dst = part + (int64_t)blockIdx.x * NN;
    for (int x = threadIdx.x; x < NN; x += blockDim.x) dst[x] = acc[x];
  }
}

// out[x] = sum_b part[b][x] (+ part[b][x^T] when SYM), fixed order.  One
// 64-lane wave per output element: lanes stride over the partials, then a
// fixed shuffle tree combines them.
template <typename T, bool SYM>
__global__ void jk_reduce_kernel(const T* __restrict__ part, int nb, int N, T* __restrict__ out) {
  const int NN = N * N;
  const int x = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (x >= NN) return;
  const int a = x / N, b = x - a * N;
  const int xt = b * N + a;
  T s = T(0);
  for (int k = lane; k < nb; k += 64) {
    s += part[(int64_t)k * NN + x];
    if (SYM) s += part[(int64_t)k * NN + xt];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) out[x] = s;
}

template <typename T>
__global__ void fill_kernel(T* __restrict__ p, int64_t n, T v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ------------------------------------------------------------- packed s8 ----
// One step of a segmented scan: the lane shifts are DPP moves (VALU), not LDS
// permutes: row_shr:k within rows of 16 lanes, then row_bcast:15 / :31 across
// rows.
template <int CTRL, int ROW_MASK, typename T>
__device__ inline void seg_step(T& a, T& b, int key) {
  const T sa = dpp_t<CTRL, ROW_MASK>(a);
  const T sb = dpp_t<CTRL, ROW_MASK>(b);
  const bool same = dpp_i<CTRL, ROW_MASK>(key, -2) == key;
  a += same ? sa : T(0);
  b += same ? sb : T(0);
}

// Inclusive prefix sums of a and b over runs of lanes with equal key (key
// nondecreasing over the valid lanes; invalid lanes carry key -1 and zeros):
// afterwards the last lane of each run holds the run's totals.
template <typename T>
__device__ inline void seg_scan2(T& a, T& b, int key) {
  seg_step<0x111, 0xf>(a, b, key);  // row_shr:1
  seg_step<0x112, 0xf>(a, b, key);  // row_shr:2
  seg_step<0x114, 0xf>(a, b, key);  // row_shr:4
  seg_step<0x118, 0xf>(a, b, key);  // row_shr:8
  seg_step<0x142, 0xa>(a, b, key);  // row_bcast:15 into rows 1, 3
  seg_step<0x143, 0xc>(a, b, key);  // row_bcast:31 into rows 2, 3
}

// (r, s) of packed lower-triangle position c = r (r + 1) / 2 + s, s <= r, for
// c < 2^24 (one correction step each way covers the 1-ulp error of sqrtf)
__device__ inline void tri_split(int c, int& r, int& s) {
  int a = (int)((sqrtf(8.0f * (float)c + 1.0f) - 1.0f) * 0.5f);
  if ((a + 1) * (a + 2) / 2 <= c) ++a;
  if (a * (a + 1) / 2 > c) --a;
  r = a;
  s = c - a * (a + 1) / 2;
}

// Contributions of one row pq = (p, q) of the packed lower triangle (rs <= pq)
// to M = J/2 - (hyb/4) K (the caller adds M + M^T), by one wavefront whose
// lanes walk rs.  Targets (r, s), (p, s), (q, s) differ between lanes; the
// targets (p, r) and (q, r) repeat over runs of consecutive rs with equal r, so
// those are summed per run with a segmented scan across the wave and added
// once per run (by the run's last lane) instead of by every lane.
//
// The row is read in groups of S8_GROUP chunks of 64 elements.  All global
// loads of a group are issued first, then all LDS reads of dm, then the LDS
// atomics: dm and M share the LDS, so a read placed after an atomic would wait
// for it, and the group's reads would become a chain of round trips.
constexpr int S8_GROUP = 4;

// Order-independent FP64 accumulation for the batched forms: every contribution
// is rounded to a multiple of 2^-JK_QBITS before it is added, so each FP64 sum
// of such values (and its intermediate sums) is exact while its magnitude stays
// below 2^(53 - JK_QBITS) = 8192 (J/K elements here are below 100).  Exact sums
// do not depend on the order of the adds: the LDS and global atomics of
// different waves and workgroups then give the same bits in every run.  The
// rounding adds at most 2^-41 (4.5e-13) per contribution.
constexpr int JK_QBITS = 40;
__device__ __forceinline__ double jk_quant(double v) { return rint(v * 0x1p40) * 0x1p-40; }

template <bool QUANT, typename TA>
__device__ __forceinline__ TA jk_term(TA v) {
  if constexpr (QUANT) return (TA)jk_quant((double)v);
  return v;
}

template <typename T, bool PRESCALED, typename TA = T, bool QUANT = false>
__device__ __forceinline__ void s8_row(const T* __restrict__ row, int pq, int p, int q, const T* __restrict__ dm,
                                       TA* __restrict__ acc, int N, T hhalf, int lane) {
  const T dpq = dm[p * N + q];
  const T fpq = (p == q && !PRESCALED) ? T(0.5) : T(1);
  const T* dmp = dm + p * N;
  const T* dmq = dm + q * N;
  T jrow = T(0);
  for (int gbase = 0; gbase <= pq; gbase += 64 * S8_GROUP) {
    T v[S8_GROUP], d_rs[S8_GROUP], d_qr[S8_GROUP], d_pr[S8_GROUP], d_qs[S8_GROUP], d_ps[S8_GROUP];
    int r[S8_GROUP], s[S8_GROUP];
#pragma unroll
    for (int u = 0; u < S8_GROUP; ++u) {
      const int rs = gbase + u * 64 + lane;
      v[u] = rs <= pq ? row[rs] : T(0);
    }
#pragma unroll
    for (int u = 0; u < S8_GROUP; ++u) {
      // positions past the row end read the row's last element's operands
      // (in bounds) and are masked below
      const int rs = min(gbase + u * 64 + lane, pq);
      tri_split(rs, r[u], s[u]);
      d_rs[u] = dm[r[u] * N + s[u]];
      d_qr[u] = dmq[r[u]];
      d_pr[u] = dmp[r[u]];
      d_qs[u] = dmq[s[u]];
      d_ps[u] = dmp[s[u]];
    }
#pragma unroll
    for (int u = 0; u < S8_GROUP; ++u) {
      const int base = gbase + u * 64;
      if (base > pq) break;  // uniform across the wavefront
      const int rs = base + lane;
      const bool valid = rs <= pq;
      // 1 / multiplicity: 2^-([p==q] + [r==s] + [pq==rs]), unless the ERIs carry it
      T w = v[u] * fpq;
      if (!PRESCALED && r[u] == s[u]) w *= T(0.5);
      if (!PRESCALED && pq == rs) w *= T(0.5);
      const T w2 = w + w;
      const T wk = -hhalf * w;
      jrow += w2 * d_rs[u];
      if (valid) {
        atomicAdd(acc + r[u] * N + s[u], jk_term<QUANT>((TA)(w2 * dpq)));
        atomicAdd(acc + p * N + s[u], jk_term<QUANT>((TA)(wk * d_qr[u])));
        atomicAdd(acc + q * N + s[u], jk_term<QUANT>((TA)(wk * d_pr[u])));
      }
      T a = wk * d_qs[u], b = wk * d_ps[u];
      const int key = valid ? r[u] : -1;
      seg_scan2(a, b, key);
      // the run of this lane's r ends here: the next position starts a new r,
      // or the chunk or the row ends
      if (valid && (s[u] == r[u] || lane == 63 || rs == pq)) {
        atomicAdd(acc + p * N + r[u], jk_term<QUANT>((TA)a));
        atomicAdd(acc + q * N + r[u], jk_term<QUANT>((TA)b));
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) jrow += __shfl_xor(jrow, off, 64);
  if (lane == 0) atomicAdd(acc + p * N + q, jk_term<QUANT>((TA)jrow));
}

// One wavefront per row pq of the packed lower triangle; lanes walk rs <= pq.
// The LDS accumulator (TA, first in the LDS, then dm) is FP64 unless
// nao^2 (sizeof(T) + 8) bytes do not fit: the waves' atomic adds come in an
// order that changes from run to run, which shows at 1e-7 of the sum in FP32
// and at 1e-16 in FP64 (ds_add_f64, no slower here than the FP32 adds).
// out != nullptr (FP64): each workgroup adds its M + M^T to the zeroed out by
// exact atomics instead of writing a partial.
template <typename T, typename TA>
__global__ void __launch_bounds__(JK_THREADS) jk_s8_kernel(const T* __restrict__ eri, int N,
                                                           const T* __restrict__ dm_g, T hhalf,
                                                           T* __restrict__ part, double* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int NN = N * N;
  const int P = N * (N + 1) / 2;
  TA* acc = reinterpret_cast<TA*>(smem);
  T* dm = reinterpret_cast<T*>(acc + NN);
  for (int x = threadIdx.x; x < NN; x += blockDim.x) {
    dm[x] = dm_g[x];
    acc[x] = TA(0);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  for (int pq = wave; pq < P; pq += nwaves) {
    int p, q;
    tri_split(pq, p, q);
    s8_row<T, false, TA>(eri + tri(pq), pq, p, q, dm, acc, N, hhalf, lane);
  }
  __syncthreads();
  if (out != nullptr) {
    for (int x = threadIdx.x; x < NN; x += blockDim.x) {
      const int i = x / N, j = x - i * N;
      atomicAdd(out + x, quant40((double)acc[x] + (double)acc[j * N + i]));
    }
    return;
  }
  T* dst = part + (int64_t)blockIdx.x * NN;
  for (int x = threadIdx.x; x < NN; x += blockDim.x) dst[x] = (T)acc[x];
}

// ------------------------------------------------------------- batched ----
// Workgroup (blockIdx.x, blockIdx.y) takes molecule order[blockIdx.x] and the
// rows pq of part blockIdx.y; converged molecules (ctl active == 0) are
// skipped.  dJK = M + M^T is written directly (one part), or added with
// atomics (several parts: the caller zeroes the output first unless ACC), or,
// with ACC, added to an FP64 accumulator (incremental build from density
// changes: the FP32 rounding then scales with |dD| instead of |D|).  The ERIs
// come divided by their multiplicity (pscf_batch_eri_scale).  LDS holds
// M (FP64) and dm of one molecule, nao^2 elements each, sized by the launch
// for the largest molecule it covers; 256 or 512 threads share them.  M is
// FP64 whatever T is: the waves' atomic adds into it come in an order that
// changes from run to run, which shows at 1e-7 of the sum in FP32 and at
// 1e-16 in FP64 (ds_add_f64, no slower here than the FP32 adds).
constexpr int JKB_MAX_THREADS = 512;

template <typename T, typename TO, bool ACC>
__global__ void __launch_bounds__(JKB_MAX_THREADS) jk_s8_batch_kernel(const BatchMol* __restrict__ mols,
                                                                      const int* __restrict__ order, int max_nao,
                                                                      const double* __restrict__ ctl,
                                                                      const T* __restrict__ eri_all,
                                                                      const T* __restrict__ dm_all, T hhalf,
                                                                      TO* __restrict__ out_all) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int m = order[blockIdx.x];
  if (ctl != nullptr && ctl[m * CTL + 1] == 0.0) return;
  const BatchMol Mo = mols[m];
  const int N = (int)Mo.nao, NN = N * N, P = N * (N + 1) / 2;
  TO* out = out_all + Mo.mat0;
  if (N > max_nao) {  // LDS was sized for max_nao: flag the block instead of overrunning it
    if (blockIdx.y == 0)
      for (int x = threadIdx.x; x < NN; x += blockDim.x) out[x] = TO(NAN);
    return;
  }
  const T* eri = eri_all + Mo.eri0;
  const T* dm_g = dm_all + Mo.mat0;
  double* acc = reinterpret_cast<double*>(smem);
  T* dm = reinterpret_cast<T*>(acc + NN);
  for (int x = threadIdx.x; x < NN; x += blockDim.x) {
    dm[x] = dm_g[x];
    acc[x] = 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int nwaves = blockDim.x >> 6;
  const int wave = (threadIdx.x >> 6) + blockIdx.y * nwaves;
  for (int pq = wave; pq < P; pq += nwaves * gridDim.y) {
    int p, q;
    tri_split(pq, p, q);
    s8_row<T, true, double, true>(eri + tri(pq), pq, p, q, dm, acc, N, hhalf, lane);
  }
  __syncthreads();
  const bool atomic = gridDim.y > 1;
  for (int x = threadIdx.x; x < NN; x += blockDim.x) {
    const int i = x / N, j = x - i * N;
    const TO v = (TO)(acc[x] + acc[j * N + i]);
    if (atomic)
      atomicAdd(out + x, v);
    else
      out[x] = ACC ? out[x] + v : v;
  }
}

// ------------------------------------------------------- batched, nao <= 64 ----
// Column form of the same contraction.  LDS float atomics cost about as much
// per lane as all the other work of an element together, so this form keeps
// the per-element targets in registers: lane s walks column s of each row
// pq = (p, q), r = 0..p in a compile-time unrolled loop, and
//   J   (r, s) += w2 dm[p][q]      -> jacc[r] of lane s, summed over every row
//                                     the wavefront handles, added to LDS once;
//   K   (p, s) += wk dm[q][r]      -> register of lane s, one LDS add per row;
//   K   (q, s) += wk dm[p][r]      -> register of lane s, one LDS add per row;
//   K   (p, r) += sum_s wk dm[q][s] \ reductions over the wavefront, 16 at a
//   K   (q, r) += sum_s wk dm[p][s] / time (8 values of r), one LDS add each;
//   J   (p, q) += sum_rs w2 dm[r][s]  (one reduction per row),
// with w the ERI divided by its multiplicity (pscf_batch_eri_scale) and
// wk = -(hyb/2) w, w2 = 2 w; the constant factors are applied per row.
__device__ inline float lane_read(float x, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l)); }
__device__ inline double lane_read(double x, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), l),
                          __builtin_amdgcn_readlane(__double2loint(x), l));
}

constexpr int JKC_MAXN = 64;
constexpr int JKC_GROUP = 8;  // r values whose ERI loads are issued together
constexpr int JKC_VWAVES = 128;  // fixed row partition of a molecule (see the kernel)

template <typename T, typename TO, bool ACC>
__global__ void __launch_bounds__(JKB_MAX_THREADS) jk_s8_cols_batch_kernel(const BatchMol* __restrict__ mols,
                                                                           const int* __restrict__ order, int max_nao,
                                                                           const double* __restrict__ ctl,
                                                                           const T* __restrict__ eri_all,
                                                                           const T* __restrict__ dm_all, T hhalf,
                                                                           TO* __restrict__ out_all) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int m = order[blockIdx.x];
  if (ctl != nullptr && ctl[m * CTL + 1] == 0.0) return;
  const BatchMol Mo = mols[m];
  const int N = (int)Mo.nao, NN = N * N, P = N * (N + 1) / 2;
  TO* out = out_all + Mo.mat0;
  if (N > max_nao || N > JKC_MAXN) {  // LDS and lanes sized for max_nao <= 64: flag the block
    if (blockIdx.y == 0)
      for (int x = threadIdx.x; x < NN; x += blockDim.x) out[x] = TO(NAN);
    return;
  }
  const T* eri = eri_all + Mo.eri0;
  const T* dm_g = dm_all + Mo.mat0;
  double* acc = reinterpret_cast<double*>(smem);
  T* dm = reinterpret_cast<T*>(acc + NN);
  for (int x = threadIdx.x; x < NN; x += blockDim.x) {
    dm[x] = dm_g[x];
    acc[x] = 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int nwaves = blockDim.x >> 6;
  // wave-uniform row index: p, q, the row pointer and dm[p][r], dm[q][r]
  // then live in scalar registers and come from scalar loads
  const int wave = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) + blockIdx.y * nwaves);
  const int sl = lane < N ? lane : N - 1;  // in-bounds column for lanes past N (their values are masked)
  // Rows go to JKC_VWAVES virtual waves (row pq to virtual wave pq mod
  // JKC_VWAVES, in ascending order), and physical wave w runs virtual waves
  // w, w + W, ... (W = waves per workgroup x workgroups per molecule).  Each
  // virtual wave's register J sums therefore cover the same rows in the same
  // order whatever W is, and their exact (quantized) adds make the result
  // independent of the launch shape.
  const int W = nwaves * gridDim.y;
  T jacc[JKC_MAXN];
  for (int vw = wave; vw < JKC_VWAVES && vw < P; vw += W) {
#pragma unroll
    for (int r = 0; r < JKC_MAXN; ++r) jacc[r] = T(0);
    for (int pq = vw; pq < P; pq += JKC_VWAVES) {
      int p, q;
      tri_split(pq, p, q);
      p = __builtin_amdgcn_readfirstlane(p);
      q = __builtin_amdgcn_readfirstlane(q);
      const T* row = eri + tri(pq);
      const T* gp = dm_g + p * N;
      const T cj = T(2) * gp[q];
      const T dp_s = lane < N ? dm[p * N + sl] : T(0);
      const T dq_s = lane < N ? dm[q * N + sl] : T(0);
      T kp = T(0), kq = T(0), jrow = T(0);
      // row pq holds (r, s) for r < p, s <= r, and for r = p, s <= q.  In FP32
      // the loads of group g + 8 are issued before group g is used (software
      // pipeline); FP64 has no registers to spare for it.
      constexpr bool PIPE = sizeof(T) == 4;
      // Loads are unconditional (a lane outside the row reads the row's first
      // element, zeroed where it is used): no branch per load, and the
      // compiler counts the loads in flight exactly.
      auto lim_of = [&](int r) { return r < p ? r : (r == p ? q : -1); };  // last valid lane of r (uniform)
      auto load_group = [&](int g0, T (&dst)[JKC_GROUP]) {
#pragma unroll
        for (int u = 0; u < JKC_GROUP; ++u) {
          const int r = g0 + u;
          dst[u] = row[lane <= lim_of(r) ? r * (r + 1) / 2 + lane : 0];
        }
      };
      T vn[JKC_GROUP];
      if constexpr (PIPE) load_group(0, vn);
#pragma unroll
      for (int g = 0; g < JKC_MAXN; g += JKC_GROUP) {
        if (g > p) break;  // uniform across the wavefront
        T v[JKC_GROUP];
        if constexpr (PIPE) {
#pragma unroll
          for (int u = 0; u < JKC_GROUP; ++u) v[u] = vn[u];
          if (g + JKC_GROUP <= p && g + JKC_GROUP < JKC_MAXN) load_group(g + JKC_GROUP, vn);
        } else {
          load_group(g, v);
        }
        // t[u] = w dm[q][s] (to K(p, g+u)), t[8+u] = w dm[p][s] (to K(q, g+u))
        T t[16];
        // r > p only in a row's last group, where w = 0: no branch per r.  dm[p][r] and dm[q][r] come from the lanes holding
        // rows p and q of dm (r is a constant of the unrolled loop).
#pragma unroll
        for (int u = 0; u < JKC_GROUP; ++u) {
          const int r = g + u;
          const T w = lane <= lim_of(r) ? v[u] : T(0);
          jacc[r] += w * cj;
          const int rr = r < N ? r : N - 1;  // rows past N: w = 0, any in-range read
          jrow += w * dm[rr * N + (lane <= rr ? lane : rr)];
          kp += w * lane_read(dq_s, r);
          kq += w * lane_read(dp_s, r);
          t[u] = w * dq_s;
          t[JKC_GROUP + u] = w * dp_s;
        }
        const T tot = transpose_reduce16(t, lane) * -hhalf;
        const int u = lane & 7;
        if (lane < 16 && g + u <= p) atomicAdd(acc + (lane < 8 ? p : q) * N + g + u, jk_quant((double)tot));
      }
      jrow = wave_sum_last(jrow) * T(2);
      if (lane < N) {
        atomicAdd(acc + p * N + lane, jk_quant((double)(kp * -hhalf)));
        atomicAdd(acc + q * N + lane, jk_quant((double)(kq * -hhalf)));
      }
      if (lane == 63) atomicAdd(acc + p * N + q, jk_quant((double)jrow));
    }
    // (predicated rather than cut short, so the loop stays unrolled and jacc in registers)
#pragma unroll
    for (int r = 0; r < JKC_MAXN; ++r)
      if (r < N && lane <= r) atomicAdd(acc + r * N + lane, jk_quant((double)jacc[r]));
  }
  __syncthreads();
  const bool atomic = gridDim.y > 1;
  for (int x = threadIdx.x; x < NN; x += blockDim.x) {
    const int i = x / N, j = x - i * N;
    const TO val = (TO)(acc[x] + acc[j * N + i]);
    if (atomic)
      atomicAdd(out + x, val);
    else
      out[x] = ACC ? out[x] + val : val;
  }
}

// Divides each molecule's packed s8 ERIs by their multiplicity
// 2^([p==q] + [r==s] + [pq==rs]) in place, the form the reference hands to its
// sparse contraction (sparse_symmetric_ERI.py:68-79, 180-182).  Workgroup
// (blockIdx.x, blockIdx.y) strides over molecule blockIdx.x's elements.
template <typename T>
__global__ void __launch_bounds__(256) eri_scale_kernel(const BatchMol* __restrict__ mols, T* __restrict__ eri_all) {
  const BatchMol Mo = mols[blockIdx.x];
  const int N = (int)Mo.nao;
  const int64_t P = (int64_t)N * (N + 1) / 2, n = P * (P + 1) / 2;
  T* eri = eri_all + Mo.eri0;
  for (int64_t e = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.y * blockDim.x) {
    int64_t a = (int64_t)((sqrt(8.0 * (double)e + 1.0) - 1.0) * 0.5);
    while (tri(a + 1) <= e) ++a;
    while (tri(a) > e) --a;
    const int pq = (int)a, rs = (int)(e - tri(a));
    int p, q, r, s;
    tri_split(pq, p, q);
    tri_split(rs, r, s);
    const int k = (p == q) + (r == s) + (pq == rs);
    eri[e] *= T(1.0 / (double)(1 << k));
  }
}

// Zeroes the dJK blocks of the active molecules order[0..gridDim.x) (before a
// split build).
template <typename TO>
__global__ void __launch_bounds__(256) zero_active_kernel(const BatchMol* __restrict__ mols,
                                                          const int* __restrict__ order,
                                                          const double* __restrict__ ctl, TO* __restrict__ out) {
  const int m = order[blockIdx.x];
  if (ctl != nullptr && ctl[m * CTL + 1] == 0.0) return;
  const BatchMol Mo = mols[m];
  const int NN = (int)(Mo.nao * Mo.nao);
  for (int x = threadIdx.x; x < NN; x += blockDim.x) out[Mo.mat0 + x] = TO(0);
}

template <typename T>
static int jk_sparse_impl(const T* eri, const int16_t* idx, int64_t nnz, const T* dm, int N, double hyb, T* out,
                          void* ws, size_t ws_bytes, hipStream_t st) {
  const int NN = N * N;
  const size_t lds = sizeof(T) * 2 * (size_t)NN;
  const T kscale = (T)(-hyb / 2.0);
  int64_t want = (nnz + JK_THREADS * 8 - 1) / (JK_THREADS * 8);
  int nb = (int)(want < 1 ? 1 : (want > JK_MAX_BLOCKS ? JK_MAX_BLOCKS : want));
  if ((int)lds <= max_lds_bytes()) {
    PSCF_REQUIRE(ws && ws_bytes >= sizeof(T) * (size_t)nb * NN, "pscf_jk_sparse: workspace too small");
    PSCF_SET_LDS((jk_sparse_kernel<T, true>), lds);
    if (nnz > 0)
      jk_sparse_kernel<T, true><<<nb, JK_THREADS, lds, st>>>(eri, idx, nnz, dm, N, kscale, (T*)ws);
    else
      fill_kernel<T><<<64, 256, 0, st>>>((T*)ws, (int64_t)NN, T(0)), nb = 1;
    PSCF_HIP(hipGetLastError());
    jk_reduce_kernel<T, false><<<(NN + 3) / 4, 256, 0, st>>>((const T*)ws, nb, N, out);
  } else {
    fill_kernel<T><<<64, 256, 0, st>>>(out, (int64_t)NN, T(0));
    if (nnz > 0) jk_sparse_kernel<T, false><<<nb, JK_THREADS, 0, st>>>(eri, idx, nnz, dm, N, kscale, out);
  }
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

// ------------------------------------------------------ block-wise direct ---
// vj, vk of ipu_einsum (pyscf_ipu/electron_repulsion/direct.py:489-504; the
// poplar_direct_s1* vertices, int2e_sph.cpp:7598-8176): the ERI blocks of the
// shell-quartet list, grouped by block size as electron_repulsion_integrals
// returns them, each block (di, dj, dk, dl) in libcint's column-major order
// with origin (i0, j0, k0, l0) and an 8-bit mask of the index permutations it
// applies (the first occurrence of each permuted origin, so every s1 element
// is counted once).  Element-parallel: thread e finds its block from the
// uniform-size classes, decodes (ci, cj, ck, cl) and adds its up to 8 J and 8 K
// contributions into LDS accumulators; per-workgroup partials are summed in a
// fixed order (reduce_parts).
constexpr int JKD_MAX_CLASSES = 8;
struct BlockClasses {
  int64_t elem0[JKD_MAX_CLASSES + 1];  // first element of class c; elem0[n] = total
  int blk0[JKD_MAX_CLASSES];           // first block of class c
  int size[JKD_MAX_CLASSES];           // elements per block of class c
  int n;
};

template <typename T>
__global__ void __launch_bounds__(JK_THREADS) jk_blocks_kernel(const T* __restrict__ eri,
                                                               const int4* __restrict__ blk, BlockClasses cls,
                                                               const T* __restrict__ dm_g, int N,
                                                               T* __restrict__ part_j, T* __restrict__ part_k) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int NN = N * N;
  T* dm = reinterpret_cast<T*>(smem);
  T* vj = dm + NN;
  T* vk = vj + NN;
  for (int x = threadIdx.x; x < NN; x += blockDim.x) {
    dm[x] = dm_g[x];
    vj[x] = T(0);
    vk[x] = T(0);
  }
  __syncthreads();
  const int64_t total = cls.elem0[cls.n];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    int c = 0;
    while (c + 1 < cls.n && e >= cls.elem0[c + 1]) ++c;
    const int size = cls.size[c];
    const int64_t rel = e - cls.elem0[c];
    const int64_t q = rel / size;
    int loc = (int)(rel - q * size);
    const int4 d = blk[cls.blk0[c] + q];
    const int di = d.x & 15, dj = (d.x >> 4) & 15, dk = (d.x >> 8) & 15;
    const int mask = (d.x >> 16) & 255;
    const int ci = loc % di;
    loc /= di;
    const int cj = loc % dj;
    loc /= dj;
    const int ck = loc % dk;
    const int cl = loc / dk;
    const int i = (d.y & 0xffff) + ci, j = (int)((unsigned)d.y >> 16) + cj;
    const int k = (d.z & 0xffff) + ck, l = (int)((unsigned)d.z >> 16) + cl;
    const T v = eri[e];
    // the eight index images (i j | k l) -> targets as the reference's vertex
    if (mask & 1) {
      atomicAdd(vj + k * N + l, dm[j * N + i] * v);
      atomicAdd(vk + i * N + l, dm[j * N + k] * v);
    }
    if (mask & 2) {
      atomicAdd(vj + l * N + k, dm[j * N + i] * v);
      atomicAdd(vk + i * N + k, dm[j * N + l] * v);
    }
    if (mask & 4) {
      atomicAdd(vj + k * N + l, dm[i * N + j] * v);
      atomicAdd(vk + j * N + l, dm[i * N + k] * v);
    }
    if (mask & 8) {
      atomicAdd(vj + l * N + k, dm[i * N + j] * v);
      atomicAdd(vk + j * N + k, dm[i * N + l] * v);
    }
    if (mask & 16) {
      atomicAdd(vj + i * N + j, dm[l * N + k] * v);
      atomicAdd(vk + k * N + j, dm[l * N + i] * v);
    }
    if (mask & 32) {
      atomicAdd(vj + j * N + i, dm[l * N + k] * v);
      atomicAdd(vk + k * N + i, dm[l * N + j] * v);
    }
    if (mask & 64) {
      atomicAdd(vj + i * N + j, dm[k * N + l] * v);
      atomicAdd(vk + l * N + j, dm[k * N + i] * v);
    }
    if (mask & 128) {
      atomicAdd(vj + j * N + i, dm[k * N + l] * v);
      atomicAdd(vk + l * N + i, dm[k * N + j] * v);
    }
  }
  __syncthreads();
  T* dj_ = part_j + (int64_t)blockIdx.x * NN;
  T* dk_ = part_k + (int64_t)blockIdx.x * NN;
  for (int x = threadIdx.x; x < NN; x += blockDim.x) {
    dj_[x] = vj[x];
    dk_[x] = vk[x];
  }
}

static int jk_blocks_nb(int64_t nelem) {
  const int64_t want = (nelem + JK_THREADS * 16 - 1) / (JK_THREADS * 16);
  return (int)(want < 1 ? 1 : (want > JK_MAX_BLOCKS ? JK_MAX_BLOCKS : want));
}

// workspace: nb vj partials, nb vk partials, then the chunk sums of reduce_parts
static size_t jk_blocks_ws_bytes(size_t es, int N, int64_t nelem) {
  const int nb = jk_blocks_nb(nelem);
  return 2 * red_align(es * (size_t)nb * N * N) + es * (size_t)red_chunks(nb) * N * N;
}

template <typename T>
static int jk_blocks_impl(const T* eri, const int4* blk, const BlockClasses& cls, const T* dm, int N, T* vj, T* vk,
                          void* ws, size_t ws_bytes, hipStream_t st) {
  const int64_t NN = (int64_t)N * N;
  const int64_t nelem = cls.elem0[cls.n];
  const int nb = jk_blocks_nb(nelem);
  const size_t lds = 3 * sizeof(T) * (size_t)NN;
  PSCF_REQUIRE((int)lds <= max_lds_bytes(), "pscf_jk_blocks: nao %d too large for the LDS accumulators", N);
  PSCF_REQUIRE(ws && ws_bytes >= jk_blocks_ws_bytes(sizeof(T), N, nelem), "pscf_jk_blocks: workspace too small");
  T* pj = (T*)ws;
  T* pk = (T*)((char*)ws + red_align(sizeof(T) * (size_t)nb * NN));
  T* lvl = (T*)((char*)ws + 2 * red_align(sizeof(T) * (size_t)nb * NN));
  PSCF_SET_LDS((jk_blocks_kernel<T>), lds);
  jk_blocks_kernel<T><<<nb, JK_THREADS, lds, st>>>(eri, blk, cls, dm, N, pj, pk);
  PSCF_HIP(hipGetLastError());
  int rc = reduce_parts<T, false>(pj, nb, N, lvl, vj, nullptr, nullptr, st);
  if (rc != PSCF_OK) return rc;
  return reduce_parts<T, false>(pk, nb, N, lvl, vk, nullptr, nullptr, st);
}

static int jk_s8_blocks(int N) {
  const int64_t P = (int64_t)N * (N + 1) / 2;
  int64_t waves = P;  // at most one wave per row
  int64_t nb = (waves + 3) / 4;
  return (int)(nb > JK_MAX_BLOCKS ? JK_MAX_BLOCKS : (nb < 1 ? 1 : nb));
}

// workspace: nb partial matrices, then the chunk sums of reduce_parts
static size_t jk_s8_ws_bytes(size_t es, int N) {
  const int nb = jk_s8_blocks(N);
  return red_align(es * (size_t)nb * N * N) + es * (size_t)red_chunks(nb) * N * N;
}

template <typename T, typename TA>
static size_t jk_s8_lds(int N) {
  return (sizeof(T) + sizeof(TA)) * (size_t)N * N;
}

template <typename T>
static int jk_s8_impl(const T* eri, int N, const T* dm, double hyb, T* out, void* ws, size_t ws_bytes,
                      hipStream_t st) {
  // FP64 accumulator when it fits, else one of T (FP32 molecules up to nao 143)
  const bool acc64 = (int)jk_s8_lds<T, double>(N) <= max_lds_bytes();
  const size_t lds = acc64 ? jk_s8_lds<T, double>(N) : jk_s8_lds<T, T>(N);
  PSCF_REQUIRE((int)lds <= max_lds_bytes(), "pscf_jk_s8: N=%d needs %zu B of LDS (max %d)", N, lds,
               max_lds_bytes());
  const int nb = jk_s8_blocks(N);
  PSCF_REQUIRE(ws && ws_bytes >= jk_s8_ws_bytes(sizeof(T), N), "pscf_jk_s8: workspace too small");
  // FP64: exact atomic sums into the zeroed output, no partials to reduce
  constexpr bool ATOMIC = sizeof(T) == 8;
  double* aout = ATOMIC ? (double*)out : nullptr;
  if (ATOMIC) {
    const int rz = zero2(aout, (int64_t)N * N, nullptr, 0, st);
    if (rz) return rz;
  }
  if (acc64) {
    PSCF_SET_LDS((jk_s8_kernel<T, double>), lds);
    jk_s8_kernel<T, double><<<nb, JK_THREADS, lds, st>>>(eri, N, dm, (T)(hyb / 2.0), (T*)ws, aout);
  } else {
    PSCF_SET_LDS((jk_s8_kernel<T, T>), lds);
    jk_s8_kernel<T, T><<<nb, JK_THREADS, lds, st>>>(eri, N, dm, (T)(hyb / 2.0), (T*)ws, aout);
  }
  PSCF_HIP(hipGetLastError());
  if (ATOMIC) return PSCF_OK;
  T* lvl = (T*)((char*)ws + red_align(sizeof(T) * (size_t)nb * N * N));
  return reduce_parts<T, true>((const T*)ws, nb, N, lvl, out, nullptr, nullptr, st);
}

// ------------------------------------------------------ allocation-free form ---
// The partial matrices and chunk sums of jk_s8_impl in a file-scope device
// array instead of a caller workspace, for callers that cannot allocate.  Each
// GPU holds its own copy (JK_SCRATCH_BYTES, until the process ends); its
// address on the current device comes from hipGetSymbolAddress once per device
// and is kept, since the array's name in host code is not a device address.
// Every partial and chunk sum is written before it is read, so nothing needs
// clearing between calls.  Calls that overlap on one GPU share the array and
// race.
constexpr size_t JK_SCRATCH_BYTES = (size_t)32 << 20;
constexpr int JK_SCRATCH_MAX_DEV = 64;
static __device__ __align__(256) unsigned char scratch[JK_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[JK_SCRATCH_MAX_DEV];

static int scratch_ptr(void** out) {
  int dev = 0;
  PSCF_HIP(hipGetDevice(&dev));
  PSCF_REQUIRE(dev >= 0 && dev < JK_SCRATCH_MAX_DEV, "scratch: device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    PSCF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PSCF_REQUIRE(p != nullptr, "scratch: no device address");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return PSCF_OK;
}

}  // namespace pscf

using namespace pscf;

extern "C" size_t pscf_jk_sparse_workspace_bytes(int dtype, int N, int64_t nnz) {
  (void)nnz;
  const size_t es = dtype == PSCF_F64 ? 8 : 4;
  return es * (size_t)JK_MAX_BLOCKS * N * N;
}

extern "C" int pscf_jk_sparse(int dtype, const void* d_eri, const int16_t* d_idx, int64_t nnz, const void* d_dm,
                              int N, double hyb, void* d_out, void* d_ws, size_t ws_bytes, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_jk_sparse: bad dtype %d", dtype);
  PSCF_REQUIRE(N > 0 && N <= 32767, "pscf_jk_sparse: N=%d out of range", N);
  PSCF_REQUIRE(nnz >= 0, "pscf_jk_sparse: nnz < 0");
  PSCF_REQUIRE(d_dm && d_out && (nnz == 0 || (d_eri && d_idx)), "pscf_jk_sparse: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64)
    return jk_sparse_impl<double>((const double*)d_eri, d_idx, nnz, (const double*)d_dm, N, hyb, (double*)d_out, d_ws,
                                  ws_bytes, st);
  return jk_sparse_impl<float>((const float*)d_eri, d_idx, nnz, (const float*)d_dm, N, hyb, (float*)d_out, d_ws,
                               ws_bytes, st);
}

extern "C" size_t pscf_jk_blocks_workspace_bytes(int dtype, int N, int64_t nelem) {
  return jk_blocks_ws_bytes(dtype == PSCF_F64 ? 8 : 4, N, nelem);
}

extern "C" int pscf_jk_blocks(int dtype, const void* d_eri, const int* d_blocks, int nclass, const int64_t* h_class_elem0,
                              const int* h_class_blk0, const int* h_class_size, const void* d_dm, int N, void* d_vj,
                              void* d_vk, void* d_ws, size_t ws_bytes, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_jk_blocks: bad dtype %d", dtype);
  PSCF_REQUIRE(N > 0 && N <= 0xffff, "pscf_jk_blocks: N=%d out of range", N);
  PSCF_REQUIRE(nclass >= 1 && nclass <= JKD_MAX_CLASSES, "pscf_jk_blocks: %d block classes (1..%d)", nclass,
               JKD_MAX_CLASSES);
  PSCF_REQUIRE(h_class_elem0 && h_class_blk0 && h_class_size && d_dm && d_vj && d_vk, "pscf_jk_blocks: null pointer");
  BlockClasses cls{};
  cls.n = nclass;
  for (int c = 0; c < nclass; ++c) {
    PSCF_REQUIRE(h_class_size[c] >= 1 && h_class_size[c] <= 15 * 15 * 15 * 15 && h_class_blk0[c] >= 0 &&
                     h_class_elem0[c + 1] >= h_class_elem0[c] && h_class_elem0[0] == 0,
                 "pscf_jk_blocks: bad class %d", c);
    PSCF_REQUIRE((h_class_elem0[c + 1] - h_class_elem0[c]) % h_class_size[c] == 0,
                 "pscf_jk_blocks: class %d is not a whole number of blocks", c);
    cls.elem0[c] = h_class_elem0[c];
    cls.blk0[c] = h_class_blk0[c];
    cls.size[c] = h_class_size[c];
  }
  cls.elem0[nclass] = h_class_elem0[nclass];
  PSCF_REQUIRE(cls.elem0[nclass] == 0 || (d_eri && d_blocks), "pscf_jk_blocks: null ERI or block table");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64)
    return jk_blocks_impl<double>((const double*)d_eri, (const int4*)d_blocks, cls, (const double*)d_dm, N,
                                  (double*)d_vj, (double*)d_vk, d_ws, ws_bytes, st);
  return jk_blocks_impl<float>((const float*)d_eri, (const int4*)d_blocks, cls, (const float*)d_dm, N, (float*)d_vj,
                               (float*)d_vk, d_ws, ws_bytes, st);
}

// ------------------------------------------------- s8 for any nao ----
// dJK = J - (hyb/2) K from the packed s8 ERIs (sparse_symmetric_ERI.py:86-137)
// for nao beyond what the LDS-resident kernels hold (dm and an N x N
// accumulator in LDS stop at nao 116 in FP32, 101 in FP64).  Only rows of dm
// and of the accumulator are in LDS; M = J/2 - (hyb/4) K accumulates in
// global memory (FP64), then dJK = M + M^T.  Every contribution added to M
// is rounded to a multiple of 2^-40 (jk_quant), so the atomic sums from
// different waves are exact and repeat bit for bit.
//   rows pass: one wave per row pq = (p, q) of the packed lower triangle,
//     lanes walking rs <= pq (as s8_row): targets (p, q) (the row's dot with
//     D), (p, s) and (q, s) (LDS atomics into the wave's copies of rows p and
//     q of M), (p, r) and (q, r) (segmented sums over runs of equal r);
//     rows p and q of D sit in the wave's LDS, D[r][s] comes from a packed
//     copy of D (index rs, coalesced).  After the row, the two row
//     accumulators and the dot go to M with global atomics.
//   columns pass: the remaining J target (r, s) += 2 w D[p][q], one wave per
//     (block of 64 rs, block of rows pq): lane rs sums over the rows in a
//     register, then adds to M once.
// w = ERI / multiplicity 2^([p==q] + [r==s] + [pq==rs]) as in s8_row.
constexpr int JK8L_WAVES = 4;      // waves per workgroup
constexpr int JK8L_ROWS = 128;     // rows pq per item of the columns pass

__global__ void __launch_bounds__(256) jk8l_prep_kernel(const double* __restrict__ dm, int N, double* __restrict__ M,
                                                        double* __restrict__ dpk) {
  const int64_t NN = (int64_t)N * N, P = tri(N);
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < NN + P; x += (int64_t)gridDim.x * blockDim.x) {
    if (x < NN) {
      M[x] = 0.0;
    } else {
      int r, s;
      tri_split((int)(x - NN), r, s);
      dpk[x - NN] = dm[r * N + s];
    }
  }
}

__device__ __forceinline__ void jk8l_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__global__ void __launch_bounds__(64 * JK8L_WAVES) jk8l_rows_kernel(const T* __restrict__ eri, int N,
                                                                    const double* __restrict__ dm,
                                                                    const double* __restrict__ dpk, double hhalf,
                                                                    double* __restrict__ M) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x & 63, wl = threadIdx.x >> 6;
  double* dp = reinterpret_cast<double*>(smem) + (size_t)4 * N * wl;  // row p of D
  double* dq = dp + N;                                                  // row q of D
  double* ap = dq + N;                                                  // row p of M (this row's part)
  double* aq = ap + N;                                                  // row q of M
  const int P = (int)tri(N);
  const int nw = gridDim.x * JK8L_WAVES;
  // longest rows first: row pq = P - 1 - item
  for (int item = __builtin_amdgcn_readfirstlane(blockIdx.x * JK8L_WAVES + wl); item < P; item += nw) {
    const int pq = P - 1 - item;
    int p, q;
    tri_split(pq, p, q);
    p = __builtin_amdgcn_readfirstlane(p);
    q = __builtin_amdgcn_readfirstlane(q);
    for (int i = lane; i < N; i += 64) {
      dp[i] = dm[(int64_t)p * N + i];
      dq[i] = dm[(int64_t)q * N + i];
      ap[i] = 0.0;
      aq[i] = 0.0;
    }
    jk8l_wave_sync();
    const T* row = eri + tri(pq);
    const double fpq = p == q ? 0.5 : 1.0;
    double jrow = 0.0;
    for (int base = 0; base <= pq; base += 64) {
      const int rs = base + lane;
      const bool valid = rs <= pq;
      const int rsc = valid ? rs : pq;
      int r, s;
      tri_split(rsc, r, s);
      const double v = valid ? (double)row[rs] : 0.0;
      double w = v * fpq;
      if (r == s) w *= 0.5;
      if (rs == pq) w *= 0.5;
      jrow += 2.0 * w * dpk[rsc];
      const double wk = -hhalf * w;
      if (valid) {
        atomicAdd(ap + s, wk * dq[r]);
        atomicAdd(aq + s, wk * dp[r]);
      }
      double a = wk * dq[s], b = wk * dp[s];
      seg_scan2(a, b, valid ? r : -1);
      if (valid && (s == r || lane == 63 || rs == pq)) {
        atomicAdd(ap + r, a);
        atomicAdd(aq + r, b);
      }
    }
    jk8l_wave_sync();
    for (int i = lane; i < N; i += 64) {
      atomicAdd(M + (int64_t)p * N + i, jk_quant(ap[i]));
      atomicAdd(M + (int64_t)q * N + i, jk_quant(aq[i]));
    }
    jrow = wave_sum_last(jrow);
    if (lane == 63) atomicAdd(M + (int64_t)p * N + q, jk_quant(jrow));
    jk8l_wave_sync();
  }
}

template <typename T>
__global__ void __launch_bounds__(256) jk8l_cols_kernel(const T* __restrict__ eri, int N,
                                                        const double* __restrict__ dpk, double* __restrict__ M) {
  const int lane = threadIdx.x & 63;
  const int P = (int)tri(N);
  const int ncb = (P + 63) / 64, nrb = (P + JK8L_ROWS - 1) / JK8L_ROWS;
  const int64_t items = (int64_t)ncb * nrb;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t it = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); it < items;
       it += nw) {
    const int cb = (int)(it / nrb), rb = (int)(it - (int64_t)cb * nrb);
    const int c0 = cb * 64;
    const int row0 = max(rb * JK8L_ROWS, c0), row1 = min((rb + 1) * JK8L_ROWS, P);
    if (row0 >= row1) continue;  // rows before the block's first column hold none of its elements
    const int rs = c0 + lane;
    int r, s;
    tri_split(rs < P ? rs : P - 1, r, s);
    const double frs = r == s ? 0.5 : 1.0;
    double jt = 0.0;
    for (int pq = row0; pq < row1; ++pq) {
      int p, q;
      tri_split(pq, p, q);
      const double v = (rs <= pq) ? (double)eri[tri(pq) + rs] : 0.0;
      double w = v * frs * (p == q ? 0.5 : 1.0);
      if (rs == pq) w *= 0.5;
      jt += 2.0 * w * dpk[pq];
    }
    if (rs < P) atomicAdd(M + (int64_t)r * N + s, jk_quant(jt));
  }
}

__global__ void __launch_bounds__(256) jk8l_sym_kernel(const double* __restrict__ M, int N, double* __restrict__ out) {
  const int64_t NN = (int64_t)N * N;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < NN; x += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(x / N), j = (int)(x - (int64_t)i * N);
    out[x] = M[x] + M[(int64_t)j * N + i];
  }
}

extern "C" size_t pscf_jk_s8_large_workspace_bytes(int N) {
  return sizeof(double) * ((size_t)N * N + (size_t)tri(N));
}

extern "C" int pscf_jk_s8_large(int dtype, const void* d_eri_s8, int N, const double* d_dm, double hyb,
                                double* d_out, void* d_ws, size_t ws_bytes, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_jk_s8_large: bad dtype %d", dtype);
  PSCF_REQUIRE(N > 0 && N <= 4096 && d_eri_s8 && d_dm && d_out, "pscf_jk_s8_large: bad arguments");
  PSCF_REQUIRE(d_ws && ws_bytes >= pscf_jk_s8_large_workspace_bytes(N), "pscf_jk_s8_large: workspace too small");
  const size_t lds = sizeof(double) * 4 * (size_t)N * JK8L_WAVES;
  PSCF_REQUIRE((int)lds <= max_lds_bytes(), "pscf_jk_s8_large: nao %d too large for the row buffers", N);
  hipStream_t st = (hipStream_t)stream;
  double* M = (double*)d_ws;
  double* dpk = M + (size_t)N * N;
  const int64_t P = tri(N);
  const int64_t prep = ((int64_t)N * N + P + 255) / 256;
  jk8l_prep_kernel<<<(unsigned)(prep < 65536 ? prep : 65536), 256, 0, st>>>(d_dm, N, M, dpk);
  PSCF_HIP(hipGetLastError());
  const int64_t rows_wg = (P + JK8L_WAVES - 1) / JK8L_WAVES;
  const unsigned g1 = (unsigned)(rows_wg < 4096 ? rows_wg : 4096);
  const int64_t items = ((P + 63) / 64) * ((P + JK8L_ROWS - 1) / JK8L_ROWS);
  const unsigned g2 = (unsigned)((items + 3) / 4 < 8192 ? (items + 3) / 4 : 8192);
  if (dtype == PSCF_F64) {
    PSCF_SET_LDS(jk8l_rows_kernel<double>, lds);
    jk8l_rows_kernel<double><<<g1, 64 * JK8L_WAVES, lds, st>>>((const double*)d_eri_s8, N, d_dm, dpk, hyb / 2.0, M);
    PSCF_HIP(hipGetLastError());
    jk8l_cols_kernel<double><<<g2, 256, 0, st>>>((const double*)d_eri_s8, N, dpk, M);
  } else {
    PSCF_SET_LDS(jk8l_rows_kernel<float>, lds);
    jk8l_rows_kernel<float><<<g1, 64 * JK8L_WAVES, lds, st>>>((const float*)d_eri_s8, N, d_dm, dpk, hyb / 2.0, M);
    PSCF_HIP(hipGetLastError());
    jk8l_cols_kernel<float><<<g2, 256, 0, st>>>((const float*)d_eri_s8, N, dpk, M);
  }
  PSCF_HIP(hipGetLastError());
  const int64_t nb = ((int64_t)N * N + 255) / 256;
  jk8l_sym_kernel<<<(unsigned)(nb < 65536 ? nb : 65536), 256, 0, st>>>(M, N, d_out);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" size_t pscf_jk_s8_workspace_bytes(int dtype, int N) {
  return jk_s8_ws_bytes(dtype == PSCF_F64 ? 8 : 4, N);
}

extern "C" int pscf_jk_s8(int dtype, const void* d_eri_s8, int N, const void* d_dm, double hyb, void* d_out,
                          void* d_ws, size_t ws_bytes, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_jk_s8: bad dtype %d", dtype);
  PSCF_REQUIRE(N > 0 && N < 65536, "pscf_jk_s8: N=%d out of range", N);
  PSCF_REQUIRE(d_eri_s8 && d_dm && d_out, "pscf_jk_s8: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64)
    return jk_s8_impl<double>((const double*)d_eri_s8, N, (const double*)d_dm, hyb, (double*)d_out, d_ws, ws_bytes, st);
  return jk_s8_impl<float>((const float*)d_eri_s8, N, (const float*)d_dm, hyb, (float*)d_out, d_ws, ws_bytes, st);
}

template <typename T, typename TO, bool ACC>
static int batch_jk_launch(dim3 grid, int threads, size_t lds, hipStream_t st, const void* mols, const int* order,
                           int max_nao, const double* ctl, const void* eri, const void* dm, double hyb, void* out) {
  if (max_nao <= JKC_MAXN) {
    PSCF_SET_LDS((jk_s8_cols_batch_kernel<T, TO, ACC>), lds);
    jk_s8_cols_batch_kernel<T, TO, ACC><<<grid, threads, lds, st>>>(
        (const BatchMol*)mols, order, max_nao, ctl, (const T*)eri, (const T*)dm, (T)(hyb / 2.0), (TO*)out);
  } else {
    PSCF_SET_LDS((jk_s8_batch_kernel<T, TO, ACC>), lds);
    jk_s8_batch_kernel<T, TO, ACC><<<grid, threads, lds, st>>>((const BatchMol*)mols, order, max_nao, ctl,
                                                               (const T*)eri, (const T*)dm, (T)(hyb / 2.0), (TO*)out);
  }
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" int pscf_batch_jk_s8(int dtype, const void* d_mols, const int* d_order, int nmol, int max_nao,
                                const double* d_ctl, const void* d_eri, const void* d_dm, double hyb, void* d_out,
                                int accumulate, int split, int threads, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_batch_jk_s8: bad dtype");
  PSCF_REQUIRE(nmol > 0 && max_nao > 0 && max_nao < 65536 && d_mols && d_order && d_eri && d_dm && d_out,
               "pscf_batch_jk_s8: bad arguments");
  PSCF_REQUIRE(split >= 1 && split <= 64, "pscf_batch_jk_s8: bad split %d", split);
  PSCF_REQUIRE(threads == 256 || threads == JKB_MAX_THREADS, "pscf_batch_jk_s8: threads %d not 256 or 512", threads);
  const size_t lds = (size_t)max_nao * max_nao * ((dtype == PSCF_F64 ? 8 : 4) + sizeof(double));
  PSCF_REQUIRE((int)lds <= max_lds_bytes(), "pscf_batch_jk_s8: nao %d too large for LDS", max_nao);
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid(nmol, split);
  if (split > 1 && !accumulate) {
    if (dtype == PSCF_F64)
      zero_active_kernel<double><<<nmol, 256, 0, st>>>((const BatchMol*)d_mols, d_order, d_ctl, (double*)d_out);
    else
      zero_active_kernel<float><<<nmol, 256, 0, st>>>((const BatchMol*)d_mols, d_order, d_ctl, (float*)d_out);
    PSCF_HIP(hipGetLastError());
  }
  if (dtype == PSCF_F64)
    return accumulate ? batch_jk_launch<double, double, true>(grid, threads, lds, st, d_mols, d_order, max_nao, d_ctl,
                                                              d_eri, d_dm, hyb, d_out)
                      : batch_jk_launch<double, double, false>(grid, threads, lds, st, d_mols, d_order, max_nao,
                                                               d_ctl, d_eri, d_dm, hyb, d_out);
  return accumulate ? batch_jk_launch<float, double, true>(grid, threads, lds, st, d_mols, d_order, max_nao, d_ctl,
                                                           d_eri, d_dm, hyb, d_out)
                    : batch_jk_launch<float, float, false>(grid, threads, lds, st, d_mols, d_order, max_nao, d_ctl,
                                                           d_eri, d_dm, hyb, d_out);
}

extern "C" int pscf_batch_eri_scale(int dtype, const void* d_mols, int nmol, int max_nao, void* d_eri, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_batch_eri_scale: bad dtype");
  PSCF_REQUIRE(nmol > 0 && max_nao > 0 && max_nao <= 4096 && d_mols && d_eri, "pscf_batch_eri_scale: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const int64_t P = (int64_t)max_nao * (max_nao + 1) / 2;
  const int64_t per = (P * (P + 1) / 2 + 255) / 256;
  const dim3 grid(nmol, (unsigned)(per < 64 ? (per < 1 ? 1 : per) : 64));
  if (dtype == PSCF_F64)
    eri_scale_kernel<double><<<grid, 256, 0, st>>>((const BatchMol*)d_mols, (double*)d_eri);
  else
    eri_scale_kernel<float><<<grid, 256, 0, st>>>((const BatchMol*)d_mols, (float*)d_eri);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

extern "C" size_t pscf_jk_s8_noalloc_max_bytes(void) { return JK_SCRATCH_BYTES; }

extern "C" int pscf_jk_s8_noalloc(int dtype, const void* d_eri_s8, int N, const void* d_dm, double hyb, void* d_out,
                                  void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_jk_s8_noalloc: bad dtype %d", dtype);
  PSCF_REQUIRE(N > 0 && N < 65536, "pscf_jk_s8_noalloc: N=%d out of range", N);
  PSCF_REQUIRE(d_eri_s8 && d_dm && d_out, "pscf_jk_s8_noalloc: null pointer");
  const size_t need = jk_s8_ws_bytes(dtype == PSCF_F64 ? 8 : 4, N);
  PSCF_REQUIRE(need <= JK_SCRATCH_BYTES, "pscf_jk_s8_noalloc: N=%d needs %zu B of scratch, above the %zu B it has", N,
               need, JK_SCRATCH_BYTES);
  void* ws = nullptr;
  int rc = scratch_ptr(&ws);
  if (rc != PSCF_OK) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == PSCF_F64)
    return jk_s8_impl<double>((const double*)d_eri_s8, N, (const double*)d_dm, hyb, (double*)d_out, ws, need, st);
  return jk_s8_impl<float>((const float*)d_eri_s8, N, (const float*)d_dm, hyb, (float*)d_out, ws, need, st);
}
