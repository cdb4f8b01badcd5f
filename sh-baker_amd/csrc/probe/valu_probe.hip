// VALU issue-rate probe: the ceiling against which the traversal and
// shading kernels (both limited by VALU instruction issue) are measured.
// Every lane runs independent dependency chains of one instruction class,
// enough waves per SIMD to hide latency, so the SIMDs issue as fast as the
// class allows. Built as its own library (libshb_probe.so); not part of the
// bake or its C ABI.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

// VALU instructions per loop iteration of each kind, as compiled for gfx950
// (checked against SQ_INSTS_VALU per dispatch).
constexpr int kBody[3] = {8, 12, 10};

// kind 0: 8 FP32 fma chains. kind 1: 4 integer chains of add, rotate, xor.
// kind 2: 4 fma chains and 2 integer chains.
template <int KIND>
__global__ void __launch_bounds__(256) k_valu_probe(float* out, unsigned long long* clk, int iters,
                                                    float s) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const bool timer = gid == 0;
  unsigned long long c0 = 0, w0 = 0;
  if (timer) {
    c0 = clock64();
    w0 = wall_clock64();
  }
  float a0 = gid * 1e-7f, a1 = a0 + 1.f, a2 = a0 + 2.f, a3 = a0 + 3.f;
  float a4 = a0 + 4.f, a5 = a0 + 5.f, a6 = a0 + 6.f, a7 = a0 + 7.f;
  uint32_t b0 = gid, b1 = gid ^ 0x9e3779b9u, b2 = gid * 3u, b3 = gid + 7u;
  uint32_t b4 = gid * 5u, b5 = gid ^ 0x85ebca6bu, b6 = gid * 11u, b7 = gid + 13u;
  for (int i = 0; i < iters; ++i) {
    if (KIND == 0 || KIND == 2) {
      a0 = fmaf(a0, s, 0.5f);
      a1 = fmaf(a1, s, 0.5f);
      a2 = fmaf(a2, s, 0.5f);
      a3 = fmaf(a3, s, 0.5f);
    }
    if (KIND == 0) {
      a4 = fmaf(a4, s, 0.5f);
      a5 = fmaf(a5, s, 0.5f);
      a6 = fmaf(a6, s, 0.5f);
      a7 = fmaf(a7, s, 0.5f);
    }
    if (KIND == 1 || KIND == 2) {
      b0 ^= b1 + 1u;
      b2 ^= b3 + 1u;
      b1 = (b1 << 7) | (b1 >> 25);
      b3 = (b3 << 7) | (b3 >> 25);
    }
    if (KIND == 1) {
      b4 ^= b5 + 1u;
      b6 ^= b7 + 1u;
      b5 = (b5 << 7) | (b5 >> 25);
      b7 = (b7 << 7) | (b7 >> 25);
    }
  }
  const float fa = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
  const uint32_t ib = b0 ^ b1 ^ b2 ^ b3 ^ b4 ^ b5 ^ b6 ^ b7;
  out[gid] = fa + (float)ib;
  if (timer) {
    clk[0] = clock64() - c0;
    clk[1] = wall_clock64() - w0;
  }
}

}  // namespace

// Compute units of device 0 (the SIMD count is 4 per CU). Returns 0 or -2.
extern "C" int shb_probe_cus(int* cus) {
  if (!cus) return -1;
  return hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, 0) == hipSuccess ? 0
                                                                                              : -2;
}

// Runs probe `kind` (0 FP32 fma, 1 integer, 2 mixed) over blocks_per_cu x
// cus blocks of 256 threads, `iters` loop iterations, `reps` timed launches
// after one untimed one. Outputs: the fastest launch in ms, the VALU wave
// instructions of the loop bodies per launch, and the shader clock of that
// launch in MHz (core-clock ticks over wall-clock ticks of thread 0).
// Returns 0, or -1 for a bad argument, -2 for a HIP error.
extern "C" int shb_probe_valu(int kind, int cus, int blocks_per_cu, int iters, int reps,
                              float* ms_out, double* wave_instr_out, double* mhz_out) {
  if (kind < 0 || kind > 2 || cus <= 0 || blocks_per_cu <= 0 || iters <= 0 || reps <= 0 ||
      !ms_out || !wave_instr_out || !mhz_out)
    return -1;
  const int blocks = cus * blocks_per_cu;
  const size_t n = (size_t)blocks * 256;
  int wall_khz = 0;
  if (hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, 0) != hipSuccess ||
      wall_khz <= 0)
    return -2;
  float* out = nullptr;
  unsigned long long* clk = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  float best = 0.f;
  double mhz = 0.0;
  if (hipMalloc(&out, n * sizeof(float)) != hipSuccess ||
      hipMalloc(&clk, 2 * sizeof(unsigned long long)) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = -2;
  } else {
    for (int r = 0; r <= reps && rc == 0; ++r) {
      if (hipEventRecord(e0, 0) != hipSuccess) rc = -2;
      if (kind == 0)
        hipLaunchKernelGGL(k_valu_probe<0>, dim3(blocks), dim3(256), 0, 0, out, clk, iters, 0.999f);
      else if (kind == 1)
        hipLaunchKernelGGL(k_valu_probe<1>, dim3(blocks), dim3(256), 0, 0, out, clk, iters, 0.999f);
      else
        hipLaunchKernelGGL(k_valu_probe<2>, dim3(blocks), dim3(256), 0, 0, out, clk, iters, 0.999f);
      float ms = 0.f;
      unsigned long long hc[2] = {0, 0};
      if (rc == 0 && (hipGetLastError() != hipSuccess || hipEventRecord(e1, 0) != hipSuccess ||
                      hipEventSynchronize(e1) != hipSuccess ||
                      hipEventElapsedTime(&ms, e0, e1) != hipSuccess ||
                      hipMemcpy(hc, clk, sizeof(hc), hipMemcpyDeviceToHost) != hipSuccess))
        rc = -2;
      if (rc == 0 && r > 0 && (best == 0.f || ms < best)) {
        best = ms;
        mhz = hc[1] ? (double)hc[0] / (double)hc[1] * (double)wall_khz * 1e-3 : 0.0;
      }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (clk) (void)hipFree(clk);
  if (out) (void)hipFree(out);
  if (rc) return rc;
  *ms_out = best;
  *wave_instr_out = (double)n / 64.0 * (double)iters * (double)kBody[kind];
  *mhz_out = mhz;
  return 0;
}
