// Sum of nb partial N x N matrices part[b][N*N] (one per workgroup of a J/K or
// XC launch), optionally plus its transpose, in two fixed-order passes:
//   pass 1  parts_sum_kernel: thread x of chunk j adds partials
//           b = 16 j .. 16 j + 15 of element x.  Consecutive threads read
//           consecutive addresses of one partial, so every partial byte is
//           fetched once and workgroups (spread over the XCDs) touch disjoint
//           memory;
//   pass 2  parts_final_kernel: out[x] = sum_j lvl[j][x], or with SYM
//           sum_j (lvl[j][x] + lvl[j][x^T]) (each pair added first, so out is
//           exactly symmetric); optionally the nb FP64 energy partials too.
#pragma once

#include "common.h"

namespace pscf {

// Order-independent FP64 sums for the single-molecule FP64 paths: each
// workgroup's partial is rounded to a multiple of 2^-40 and added to the
// output with an atomic.  Sums of such values are exact in FP64 while they
// stay below 2^13 (Fock-matrix elements and E_xc here are far below), so the
// result does not depend on the order of the atomics and repeats bit for
// bit; the rounding adds at most 2^-41 per workgroup and element.
__device__ __forceinline__ double quant40(double v) { return rint(v * 0x1p40) * 0x1p-40; }

// Zeroes a[0..na) and b[0..nb) (the outputs of an atomic sum).
static __global__ void __launch_bounds__(256) zero2_kernel(double* __restrict__ a, int64_t na, double* __restrict__ b,
                                                            int nb) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < na + nb; x += (int64_t)gridDim.x * blockDim.x) {
    if (x < na)
      a[x] = 0.0;
    else
      b[x - na] = 0.0;
  }
}

inline int zero2(double* a, int64_t na, double* b, int nb, hipStream_t st) {
  const int64_t blocks = (na + nb + 255) / 256;
  zero2_kernel<<<(unsigned)(blocks < 1 ? 1 : (blocks > 1024 ? 1024 : blocks)), 256, 0, st>>>(a, na, b, nb);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

constexpr int RED_CHUNK = 16;
constexpr int RED_THREADS = 256;

__host__ __device__ inline int red_chunks(int nb) { return (nb + RED_CHUNK - 1) / RED_CHUNK; }

__host__ __device__ inline size_t red_align(size_t bytes) { return (bytes + 255) & ~(size_t)255; }

template <typename T>
__global__ void __launch_bounds__(RED_THREADS) parts_sum_kernel(const T* __restrict__ part, int nb, int NN,
                                                                  T* __restrict__ lvl) {
  const int x = blockIdx.x * RED_THREADS + threadIdx.x;
  if (x >= NN) return;
  const int b0 = blockIdx.y * RED_CHUNK;
  const int n = nb - b0 < RED_CHUNK ? nb - b0 : RED_CHUNK;  // >= 1
  const T* p = part + (int64_t)b0 * NN + x;
  T v[RED_CHUNK];
#pragma unroll
  for (int b = 0; b < RED_CHUNK; ++b) {
    const T t = p[(int64_t)(b < n ? b : n - 1) * NN];  // clamped: every load in bounds
    v[b] = b < n ? t : T(0);
  }
#pragma unroll
  for (int w = RED_CHUNK / 2; w > 0; w >>= 1)
#pragma unroll
    for (int b = 0; b < w; ++b) v[b] += v[b + w];
  lvl[(int64_t)blockIdx.y * NN + x] = v[0];
}

template <typename T, bool SYM>
__global__ void __launch_bounds__(RED_THREADS) parts_final_kernel(const T* __restrict__ lvl, int nc, int N,
                                                                    T* __restrict__ out,
                                                                    const double* __restrict__ part_e, int nb,
                                                                    double* __restrict__ exc) {
  const int NN = N * N;
  const int x = blockIdx.x * RED_THREADS + threadIdx.x;
  if (x < NN) {
    const int i = x / N, j = x - i * N;
    const int xt = j * N + i;
    T a[RED_CHUNK], b[RED_CHUNK];  // every load issued before the sums (nc <= RED_CHUNK)
#pragma unroll
    for (int c = 0; c < RED_CHUNK; ++c) {
      const int64_t o = (int64_t)(c < nc ? c : nc - 1) * NN;
      a[c] = lvl[o + x];
      b[c] = SYM ? lvl[o + xt] : T(0);
    }
    T s = T(0);
#pragma unroll
    for (int c = 0; c < RED_CHUNK; ++c)
      if (c < nc) s += SYM ? a[c] + b[c] : a[c];
    out[x] = s;
  }
  // energy partials: lanes of wave 0 of block 0 (Neumaier each), fixed tree
  if (exc != nullptr && blockIdx.x == 0 && threadIdx.x < 64) {
    double s = 0.0, c = 0.0;
    for (int b = threadIdx.x; b < nb; b += 64) {
      const double v = part_e[b], t = s + v;
      c += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
      s = t;
    }
    s += c;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (threadIdx.x == 0) *exc = s;
  }
}

// Launches both passes on st; lvl holds red_chunks(nb) * N * N elements.
// nb <= RED_CHUNK^2 = 256 partials.
template <typename T, bool SYM>
inline int reduce_parts(const T* part, int nb, int N, T* lvl, T* out, const double* part_e, double* exc,
                        hipStream_t st) {
  const int NN = N * N, nc = red_chunks(nb);
  PSCF_REQUIRE(nb >= 1 && nc <= RED_CHUNK, "reduce_parts: %d partials (1..%d)", nb, RED_CHUNK * RED_CHUNK);
  const int gx = (NN + RED_THREADS - 1) / RED_THREADS;
  parts_sum_kernel<T><<<dim3(gx, nc), RED_THREADS, 0, st>>>(part, nb, NN, lvl);
  PSCF_HIP(hipGetLastError());
  parts_final_kernel<T, SYM><<<gx, RED_THREADS, 0, st>>>(lvl, nc, N, out, part_e, nb, exc);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}

}  // namespace pscf
