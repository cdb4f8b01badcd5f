// Exhaustive check of the light loop's corrected square root and reciprocal
// (sqrt_core, rcp_core in shb_common.h) against sqrtf and 1.0f / x, the
// correctly rounded IEEE operations, over every float of the ranges
// inv_dist uses them on. Built into libshb_probe.so; not part of the bake or
// its C ABI.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../shb_common.h"

namespace {

// Every bit pattern lo .. lo + n - 1 (positive floats, so consecutive
// patterns are consecutive values); mismatch count and the lowest
// mismatching pattern per operation (0: sqrt, 1: reciprocal).
template <int OP>
__global__ void __launch_bounds__(256) k_rounding(uint32_t lo, uint32_t n,
                                                  unsigned long long* bad, uint32_t* first) {
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t local = 0, low = 0xffffffffu;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t bits = lo + i;
    const float x = shb::as_float(bits);
    float a, b;
    if (OP == 0) {
      a = shb::sqrt_core(x);
      b = sqrtf(x);
    } else {
      a = shb::rcp_core(x);
      b = 1.0f / x;
    }
    if (shb::as_uint(a) != shb::as_uint(b)) {
      ++local;
      low = bits < low ? bits : low;
    }
    if (i > 0xffffffffu - stride) break;  // the next index would wrap
  }
  if (local) {
    atomicAdd(bad + OP, (unsigned long long)local);
    atomicMin(first + OP, low);
  }
}

}  // namespace

// Checks sqrt_core over [2^-96, 2^126] and rcp_core over [2^-48, 2^63]
// (every float of each). out[0], out[1]: sqrt mismatches and floats
// checked; out[2], out[3]: the same for the reciprocal; out[4], out[5]: the
// lowest mismatching bit pattern of each (0xffffffff: none). Returns 0, -1
// for a null argument, -2 for a HIP error.
extern "C" int shb_probe_rounding(unsigned long long* out) {
  if (!out) return -1;
  unsigned long long* bad = nullptr;
  uint32_t* first = nullptr;
  int rc = 0;
  const uint32_t s_lo = shb::as_uint(0x1p-96f), s_hi = shb::as_uint(0x1p126f);
  const uint32_t r_lo = shb::as_uint(0x1p-48f), r_hi = shb::as_uint(0x1p63f);
  const uint32_t s_n = s_hi - s_lo + 1u, r_n = r_hi - r_lo + 1u;
  unsigned long long hb[2] = {0, 0};
  uint32_t hf[2] = {0xffffffffu, 0xffffffffu};
  if (hipMalloc(&bad, sizeof(hb)) != hipSuccess || hipMalloc(&first, sizeof(hf)) != hipSuccess ||
      hipMemcpy(bad, hb, sizeof(hb), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(first, hf, sizeof(hf), hipMemcpyHostToDevice) != hipSuccess) {
    rc = -2;
  } else {
    hipLaunchKernelGGL(k_rounding<0>, dim3(8192), dim3(256), 0, 0, s_lo, s_n, bad, first);
    hipLaunchKernelGGL(k_rounding<1>, dim3(8192), dim3(256), 0, 0, r_lo, r_n, bad, first);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(hb, bad, sizeof(hb), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hf, first, sizeof(hf), hipMemcpyDeviceToHost) != hipSuccess)
      rc = -2;
  }
  if (bad) (void)hipFree(bad);
  if (first) (void)hipFree(first);
  out[0] = hb[0];
  out[1] = s_n;
  out[2] = hb[1];
  out[3] = r_n;
  out[4] = hf[0];
  out[5] = hf[1];
  return rc;
}
