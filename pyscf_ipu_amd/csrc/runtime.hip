// Error plumbing and small helpers shared by all entry points of libpyscf_mi355x.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "common.h"

namespace pscf {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int device_max_lds() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 65536;
  if (cached[dev] == 0) {
    int v = 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) v = 65536;
    cached[dev] = v;
  }
  return cached[dev];
}

int check_hip(hipError_t e, const char* what) {
  set_error("%s: %s", what, hipGetErrorString(e));
  return PSCF_EHIP;
}

template <typename S, typename D>
__global__ void cast_kernel(const S* __restrict__ src, D* __restrict__ dst, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (D)src[i];
}

}  // namespace pscf

using namespace pscf;

extern "C" const char* pscf_last_error(void) { return g_err; }

extern "C" int pscf_abi_version(void) { return 1; }

extern "C" int pscf_cast(int src_dtype, const void* d_src, int dst_dtype, void* d_dst, int64_t n, void* stream) {
  PSCF_REQUIRE(n >= 0, "pscf_cast: n < 0");
  PSCF_REQUIRE((src_dtype == PSCF_F64 || src_dtype == PSCF_F32) && (dst_dtype == PSCF_F64 || dst_dtype == PSCF_F32),
               "pscf_cast: bad dtype");
  if (n == 0) return PSCF_OK;
  hipStream_t s = (hipStream_t)stream;
  int blocks = (int)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
  if (src_dtype == PSCF_F64 && dst_dtype == PSCF_F32)
    cast_kernel<<<blocks, 256, 0, s>>>((const double*)d_src, (float*)d_dst, n);
  else if (src_dtype == PSCF_F32 && dst_dtype == PSCF_F64)
    cast_kernel<<<blocks, 256, 0, s>>>((const float*)d_src, (double*)d_dst, n);
  else if (src_dtype == PSCF_F64)
    cast_kernel<<<blocks, 256, 0, s>>>((const double*)d_src, (double*)d_dst, n);
  else
    cast_kernel<<<blocks, 256, 0, s>>>((const float*)d_src, (float*)d_dst, n);
  PSCF_HIP(hipGetLastError());
  return PSCF_OK;
}
