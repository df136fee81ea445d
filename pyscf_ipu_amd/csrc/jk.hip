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
#include "batch.h"
#include "common.h"

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
// Contributions of one row pq = (p, q) of the packed lower triangle (rs <= pq)
// to M = J/2 - (hyb/4) K (the caller adds M + M^T), by one wavefront whose
// lanes walk rs.  Targets (r, s), (p, s), (q, s) differ between lanes; the
// targets (p, r) and (q, r) repeat over runs of consecutive rs with equal r, so
// those are summed per run with a segmented suffix scan across the wave and
// added once per run instead of by every lane.
constexpr int S8_GROUP = 4;

template <typename T>
__device__ inline void s8_row(const T* __restrict__ row, int pq, int p, int q, const T* __restrict__ dm,
                              T* __restrict__ acc, const unsigned* __restrict__ pair, int N, T hhalf, int lane) {
  const T dpq = dm[p * N + q];
  T jrow = T(0);
  // the row is read in groups of S8_GROUP chunks of 64 elements whose global
  // loads are all issued before any is used, so their latency overlaps
  for (int gbase = 0; gbase <= pq; gbase += 64 * S8_GROUP) {
    T vv[S8_GROUP];
#pragma unroll
    for (int u = 0; u < S8_GROUP; ++u) {
      const int rs = gbase + u * 64 + lane;
      vv[u] = rs <= pq ? row[rs] : T(0);
    }
#pragma unroll
    for (int u = 0; u < S8_GROUP; ++u) {
      const int base = gbase + u * 64;
      if (base > pq) continue;  // uniform across the wavefront
      const int rs = base + lane;
      const bool valid = rs <= pq;
      int r = -1, s = 0;
      T a = T(0), b = T(0);
      if (valid) {
        const unsigned rr = pair[rs];
        r = rr & 0xffff;
        s = rr >> 16;
        // 1 / multiplicity: 2^-([p==q] + [r==s] + [pq==rs])
        T v = vv[u];
        if (p == q) v *= T(0.5);
        if (r == s) v *= T(0.5);
        if (pq == rs) v *= T(0.5);
        const T v2 = v + v;
        atomicAdd(acc + r * N + s, v2 * dpq);
        jrow += v2 * dm[r * N + s];
        const T vk = -hhalf * v;
        atomicAdd(acc + p * N + s, vk * dm[q * N + r]);
        atomicAdd(acc + q * N + s, vk * dm[p * N + r]);
        a = vk * dm[q * N + s];
        b = vk * dm[p * N + s];
      }
      // segmented suffix sums over lanes of equal r (r is nondecreasing in rs)
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const T ta = __shfl_down(a, off, 64);
        const T tb = __shfl_down(b, off, 64);
        const int tr = __shfl_down(r, off, 64);
        if (lane + off < 64 && tr == r) {
          a += ta;
          b += tb;
        }
      }
      const int rprev = __shfl_up(r, 1, 64);
      if (valid && (lane == 0 || rprev != r)) {
        atomicAdd(acc + p * N + r, a);
        atomicAdd(acc + q * N + r, b);
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) jrow += __shfl_xor(jrow, off, 64);
  if (lane == 0) atomicAdd(acc + p * N + q, jrow);
}

// One wavefront per row pq of the packed lower triangle; lanes walk rs <= pq.
template <typename T>
__global__ void __launch_bounds__(JK_THREADS) jk_s8_kernel(const T* __restrict__ eri, int N,
                                                           const T* __restrict__ dm_g, T hhalf,
                                                           T* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int NN = N * N;
  const int P = N * (N + 1) / 2;
  T* dm = reinterpret_cast<T*>(smem);
  T* acc = dm + NN;
  unsigned* pair = reinterpret_cast<unsigned*>(acc + NN);
  for (int x = threadIdx.x; x < NN; x += blockDim.x) {
    dm[x] = dm_g[x];
    acc[x] = T(0);
  }
  for (int c = threadIdx.x; c < P; c += blockDim.x) {
    int a = (int)((sqrtf(8.0f * (float)c + 1.0f) - 1.0f) * 0.5f);
    while (tri(a + 1) <= c) ++a;
    while (tri(a) > c) --a;
    pair[c] = (unsigned)a | ((unsigned)(c - tri(a)) << 16);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  for (int pq = wave; pq < P; pq += nwaves) {
    const unsigned pp = pair[pq];
    s8_row<T>(eri + tri(pq), pq, pp & 0xffff, pp >> 16, dm, acc, pair, N, hhalf, lane);
  }
  __syncthreads();
  T* dst = part + (int64_t)blockIdx.x * NN;
  for (int x = threadIdx.x; x < NN; x += blockDim.x) dst[x] = acc[x];
}

// ------------------------------------------------------------- batched ----
// Workgroup (blockIdx.x, blockIdx.y) takes molecule order[blockIdx.x] and the
// rows pq of part blockIdx.y; converged molecules (ctl active == 0) are
// skipped.  dJK = M + M^T is written directly (one part), or added with
// atomics (several parts: the caller zeroes the output first unless ACC), or,
// with ACC, added to an FP64 accumulator (incremental build from density
// changes: the FP32 rounding then scales with |dD| instead of |D|).
template <typename T, typename TO, bool ACC>
__global__ void __launch_bounds__(JK_THREADS) jk_s8_batch_kernel(const BatchMol* __restrict__ mols,
                                                                 const int* __restrict__ order,
                                                                 const double* __restrict__ ctl,
                                                                 const T* __restrict__ eri_all,
                                                                 const T* __restrict__ dm_all, T hhalf,
                                                                 TO* __restrict__ out_all) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int m = order[blockIdx.x];
  if (ctl != nullptr && ctl[m * CTL + 1] == 0.0) return;
  const BatchMol Mo = mols[m];
  const int N = (int)Mo.nao, NN = N * N, P = N * (N + 1) / 2;
  const T* eri = eri_all + Mo.eri0;
  const T* dm_g = dm_all + Mo.mat0;
  T* dm = reinterpret_cast<T*>(smem);
  T* acc = dm + NN;
  unsigned* pair = reinterpret_cast<unsigned*>(acc + NN);
  for (int x = threadIdx.x; x < NN; x += blockDim.x) {
    dm[x] = dm_g[x];
    acc[x] = T(0);
  }
  for (int c = threadIdx.x; c < P; c += blockDim.x) {
    int a = (int)((sqrtf(8.0f * (float)c + 1.0f) - 1.0f) * 0.5f);
    while (tri(a + 1) <= c) ++a;
    while (tri(a) > c) --a;
    pair[c] = (unsigned)a | ((unsigned)(c - tri(a)) << 16);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int nwaves = blockDim.x >> 6;
  const int wave = (threadIdx.x >> 6) + blockIdx.y * nwaves;
  for (int pq = wave; pq < P; pq += nwaves * gridDim.y) {
    const unsigned pp = pair[pq];
    s8_row<T>(eri + tri(pq), pq, pp & 0xffff, pp >> 16, dm, acc, pair, N, hhalf, lane);
  }
  __syncthreads();
  TO* out = out_all + Mo.mat0;
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

// Zeroes the dJK blocks of the active molecules (before a split build).
template <typename TO>
__global__ void __launch_bounds__(256) zero_active_kernel(const BatchMol* __restrict__ mols,
                                                          const double* __restrict__ ctl, TO* __restrict__ out) {
  const int m = blockIdx.x;
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

static int jk_s8_blocks(int N) {
  const int64_t P = (int64_t)N * (N + 1) / 2;
  int64_t waves = P;  // at most one wave per row
  int64_t nb = (waves + 3) / 4;
  return (int)(nb > JK_MAX_BLOCKS ? JK_MAX_BLOCKS : (nb < 1 ? 1 : nb));
}

template <typename T>
static size_t jk_s8_lds(int N) {
  return sizeof(T) * 2 * (size_t)N * N + sizeof(unsigned) * (size_t)N * (N + 1) / 2;
}

template <typename T>
static int jk_s8_impl(const T* eri, int N, const T* dm, double hyb, T* out, void* ws, size_t ws_bytes,
                      hipStream_t st) {
  const size_t lds = jk_s8_lds<T>(N);
  PSCF_REQUIRE((int)lds <= max_lds_bytes(), "pscf_jk_s8: N=%d needs %zu B of LDS (max %d)", N, lds,
               max_lds_bytes());
  const int nb = jk_s8_blocks(N);
  PSCF_REQUIRE(ws && ws_bytes >= sizeof(T) * (size_t)nb * N * N, "pscf_jk_s8: workspace too small");
  PSCF_SET_LDS(jk_s8_kernel<T>, lds);
  jk_s8_kernel<T><<<nb, JK_THREADS, lds, st>>>(eri, N, dm, (T)(hyb / 2.0), (T*)ws);
  PSCF_HIP(hipGetLastError());
  jk_reduce_kernel<T, true><<<(N * N + 3) / 4, 256, 0, st>>>((const T*)ws, nb, N, out);
  PSCF_HIP(hipGetLastError());
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

extern "C" size_t pscf_jk_s8_workspace_bytes(int dtype, int N) {
  const size_t es = dtype == PSCF_F64 ? 8 : 4;
  return es * (size_t)jk_s8_blocks(N) * N * N;
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

extern "C" int pscf_batch_jk_s8(int dtype, const void* d_mols, const int* d_order, int nmol, int max_nao,
                                const double* d_ctl, const void* d_eri, const void* d_dm, double hyb, void* d_out,
                                int accumulate, int split, void* stream) {
  PSCF_REQUIRE(dtype == PSCF_F64 || dtype == PSCF_F32, "pscf_batch_jk_s8: bad dtype");
  PSCF_REQUIRE(nmol > 0 && max_nao > 0 && max_nao < 65536 && d_mols && d_order && d_eri && d_dm && d_out,
               "pscf_batch_jk_s8: bad arguments");
  PSCF_REQUIRE(split >= 1 && split <= 64, "pscf_batch_jk_s8: bad split %d", split);
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid(nmol, split);
  if (split > 1 && !accumulate) {
    if (dtype == PSCF_F64)
      zero_active_kernel<double><<<nmol, 256, 0, st>>>((const BatchMol*)d_mols, d_ctl, (double*)d_out);
    else
      zero_active_kernel<float><<<nmol, 256, 0, st>>>((const BatchMol*)d_mols, d_ctl, (float*)d_out);
    PSCF_HIP(hipGetLastError());
  }
  if (dtype == PSCF_F64) {
    const size_t lds = jk_s8_lds<double>(max_nao);
    PSCF_REQUIRE((int)lds <= max_lds_bytes(), "pscf_batch_jk_s8: nao %d too large for LDS", max_nao);
    if (accumulate) {
      PSCF_SET_LDS((jk_s8_batch_kernel<double, double, true>), lds);
      jk_s8_batch_kernel<double, double, true><<<grid, JK_THREADS, lds, st>>>(
          (const BatchMol*)d_mols, d_order, d_ctl, (const double*)d_eri, (const double*)d_dm, hyb / 2.0,
          (double*)d_out);
    } else {
      PSCF_SET_LDS((jk_s8_batch_kernel<double, double, false>), lds);
      jk_s8_batch_kernel<double, double, false><<<grid, JK_THREADS, lds, st>>>(
          (const BatchMol*)d_mols, d_order, d_ctl, (const double*)d_eri, (const double*)d_dm, hyb / 2.0,
          (double*)d_out);
    }
  } else {
    const size_t lds = jk_s8_lds<float>(max_nao);
    PSCF_REQUIRE((int)lds <= max_lds_bytes(), "pscf_batch_jk_s8: nao %d too large for LDS", max_nao);
    if (accumulate) {
      PSCF_SET_LDS((jk_s8_batch_kernel<float, double, true>), lds);
      jk_s8_batch_kernel<float, double, true><<<grid, JK_THREADS, lds, st>>>(
          (const BatchMol*)d_mols, d_order, d_ctl, (const float*)d_eri, (const float*)d_dm, (float)(hyb / 2.0),
          (double*)d_out);
    } else {
      PSCF_SET_LDS((jk_s8_batch_kernel<float, float, false>), lds);
      jk_s8_batch_kernel<float, float, false><<<grid, JK_THREADS, lds, st>>>(
          (const BatchMol*)d_mols, d_order, d_ctl, (const float*)d_eri, (const float*)d_dm, (float)(hyb / 2.0),
          (float*)d_out);
    }
  }
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}
