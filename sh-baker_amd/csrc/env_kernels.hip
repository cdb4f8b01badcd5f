// Environment SH: projection of an equirectangular (LatLong) radiance map onto
// the nine real SH basis functions, src/scene.cpp:31-71 (ProjectLatLongMap).
//
// Per texel (x, y) of a w x h map: theta = (y + 0.5) pi / h from the zenith
// (+Y), phi = (x + 0.5) 2 pi / w from +Z towards +X, direction
// d = (sin t sin p, cos t, sin t cos p), weight d_omega = (2 pi / w)(pi / h)
// sin t; rows with sin t < 1e-4 are skipped, as in the reference.
//
// Separable form. In a column, sin p and cos p are constants, so each basis
// function (src/sh_coeffs.cpp:50-78) is a column factor times one of six row
// functions R_j = {1, cos t, sin t, sin t cos t, sin^2 t, cos^2 t}:
//   Y0 = k0 R0, Y1 = k1 R1, Y2 = k1 cos p R2, Y3 = k1 sin p R2,
//   Y4 = k21 sin p R3, Y5 = k21 cos p R3, Y6 = 3 k20 cos^2 p R4 - k20 R0,
//   Y7 = k21 sin p cos p R4, Y8 = k22 sin^2 p R4 - k22 R5.
// A thread owns one column of a block's row strip and sums S_jc = sum_y
// L_c(x, y) d_omega_y R_j(y) (18 FMAs per texel against a per-row LDS table
// of d_omega R_j), then applies the column factors once, in FP64. The sums
// are FP32 over 8 rows, then FP64 per thread, per block and across blocks
// in a fixed order: deterministic, and closer to the exact sum of the
// reference's terms than its serial float sum (the terms themselves agree
// with the reference's float products to rounding).
//
// HBM-bound: each texel's three floats are read once; loads of 8 rows are
// issued before their first use.
#include <algorithm>
#include <mutex>

#include "kernels.h"

namespace shb {

namespace {

constexpr float kEnvPi = 3.14159265359f;     // scene.cpp:12
constexpr float kEnvTwoPi = 6.28318530718f;  // scene.cpp:13
constexpr double kY0 = 0.282095, kY1 = 0.488603, kY21 = 1.092548, kY20 = 0.315392,
                 kY22 = 0.546274;  // sh_coeffs.cpp:27-31
// One wave per block: its 64 columns x up to 256 rows. About 2048 waves
// balance the per-wave fixed cost (row table, final reductions) against
// the number of waves in flight (measured on MI355X, 8192 x 4096 map: 2048
// waves of 256 rows 80 us; 4096 of 128 rows 107 us; 1024 of 512 rows 135 us).
constexpr int kEnvThreads = 64;
constexpr int kEnvWaves = kEnvThreads / 64;
constexpr int kEnvMaxRows = 256;  // rows per block (LDS table size)
constexpr int kEnvBatch = 8;      // rows loaded together, summed in FP32, then added in FP64
constexpr int kEnvBlocks = 2048;  // target grid size

}  // namespace

// One block: kEnvThreads columns x rows_per_block rows. Writes its 27 sums to
// partials[i * nblocks + block] (coefficient-major, for k_env_finish).
__global__ void __launch_bounds__(kEnvThreads)
    k_env_latlong(const float* __restrict__ px, int w, int h, int ch, int rows_per_block,
                  double* __restrict__ partials) {
  __shared__ float4 row_a[kEnvMaxRows];  // d_omega * (1, cos t, sin t, sin t cos t)
  __shared__ float2 row_b[kEnvMaxRows];  // d_omega * (sin^2 t, cos^2 t); d_omega 0: skipped
  __shared__ double red[kEnvWaves][27];
  const float d_phi = kEnvTwoPi / (float)w;
  const float d_theta = kEnvPi / (float)h;
  const int y0 = blockIdx.y * rows_per_block;
  const int nrows = min(rows_per_block, h - y0);
  for (int r = threadIdx.x; r < nrows; r += kEnvThreads) {
    const float theta = ((float)(y0 + r) + 0.5f) * d_theta;
    const float st = sinf(theta), ct = cosf(theta);
    const float dw = st < 1e-4f ? 0.0f : d_phi * d_theta * st;
    row_a[r] = make_float4(dw, dw * ct, dw * st, dw * (st * ct));
    row_b[r] = make_float2(dw * (st * st), dw * (ct * ct));
  }
  __syncthreads();
  double acc[18];  // S_jc, j-major
#pragma unroll
  for (int i = 0; i < 18; ++i) acc[i] = 0.0;
  const int x = blockIdx.x * kEnvThreads + threadIdx.x;
  if (x < w) {
    const float* p = px + ((size_t)y0 * w + x) * ch;
    const size_t row_stride = (size_t)w * ch;
    for (int r0 = 0; r0 < nrows; r0 += kEnvBatch) {
      float v[kEnvBatch][3];
#pragma unroll
      for (int k = 0; k < kEnvBatch; ++k) {
        if (r0 + k < nrows) {
          const float* q = p + (size_t)(r0 + k) * row_stride;
          v[k][0] = q[0];
          v[k][1] = q[1];
          v[k][2] = q[2];
        }
      }
      float part[18];
#pragma unroll
      for (int i = 0; i < 18; ++i) part[i] = 0.0f;
#pragma unroll
      for (int k = 0; k < kEnvBatch; ++k) {
        // the row is the same for the whole block: uniform branches
        if (r0 + k >= nrows) break;
        const float4 a = row_a[r0 + k];
        if (a.x == 0.0f) continue;  // skipped row: its texels are not read into the sum
        const float2 b = row_b[r0 + k];
        const float R[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
#pragma unroll
        for (int j = 0; j < 6; ++j)
#pragma unroll
          for (int c = 0; c < 3; ++c) part[3 * j + c] = fmaf(v[k][c], R[j], part[3 * j + c]);
      }
#pragma unroll
      for (int i = 0; i < 18; ++i) acc[i] += (double)part[i];
    }
  }
  // column factors (FP64), then wave sums by shuffles and the four waves'
  // sums in order (one barrier)
  const float phi = ((float)min(x, w - 1) + 0.5f) * d_phi;
  const double sp = (double)sinf(phi), cp = (double)cosf(phi);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    const int k = i / 3, c = i % 3;
    const double* S = acc;  // S_j at acc + 3 j
    double s;
    switch (k) {
      case 0: s = kY0 * S[c]; break;
      case 1: s = kY1 * S[3 + c]; break;
      case 2: s = kY1 * cp * S[6 + c]; break;
      case 3: s = kY1 * sp * S[6 + c]; break;
      case 4: s = kY21 * sp * S[9 + c]; break;
      case 5: s = kY21 * cp * S[9 + c]; break;
      case 6: s = 3.0 * kY20 * cp * cp * S[12 + c] - kY20 * S[c]; break;
      case 7: s = kY21 * sp * cp * S[12 + c]; break;
      default: s = kY22 * sp * sp * S[12 + c] - kY22 * S[15 + c]; break;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) red[wave][i] = s;
  }
  __syncthreads();
  if (threadIdx.x < 27) {
    double s = 0.0;
    for (int k = 0; k < kEnvWaves; ++k) s += red[k][threadIdx.x];
    const size_t nblocks = (size_t)gridDim.x * gridDim.y;
    partials[threadIdx.x * nblocks + (size_t)blockIdx.y * gridDim.x + blockIdx.x] = s;
  }
}

// Block i sums partials[i * nblocks + b] over b in a fixed order into out[i]
// (FP64) and / or out_f[i] (rounded to float); either may be null.
constexpr int kFinishThreads = 256;
__global__ void __launch_bounds__(kFinishThreads)
    k_env_finish(const double* __restrict__ partials, int nblocks, double* __restrict__ out,
                 float* __restrict__ out_f) {
  __shared__ double red[kFinishThreads / 64];
  const double* p = partials + (size_t)blockIdx.x * nblocks;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += kFinishThreads) s += p[b];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < kFinishThreads / 64; ++k) t += red[k];
    if (out) out[blockIdx.x] = t;
    if (out_f) out_f[blockIdx.x] = (float)t;
  }
}

EnvGrid env_grid(int w, int h) {
  EnvGrid g;
  g.gx = (w + kEnvThreads - 1) / kEnvThreads;
  // about kEnvBlocks blocks, each a strip of at most kEnvMaxRows rows
  long long want = ((long long)h * g.gx + kEnvBlocks - 1) / kEnvBlocks;
  g.rows_per_block = (int)std::min<long long>(kEnvMaxRows, std::max<long long>(1, want));
  g.gy = (h + g.rows_per_block - 1) / g.rows_per_block;
  return g;
}

hipError_t launch_env_latlong(const float* px, int w, int h, int ch, const EnvGrid& g,
                              double* partials, double* out, float* out_f, hipStream_t st) {
  hipLaunchKernelGGL(k_env_latlong, dim3(g.gx, g.gy), dim3(kEnvThreads), 0, st, px, w, h, ch,
                     g.rows_per_block, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_env_finish, dim3(27), dim3(kFinishThreads), 0, st, partials, g.gx * g.gy,
                     out, out_f);
  return hipGetLastError();
}

// ------------------------------------------------ fixed scratch (no allocation)
// The per-block partial sums of shb_env_sh_latlong_noalloc live in this
// array, one copy per GPU for the life of the process. In host code its name
// is only a stand-in: kernels get the device address that
// hipGetSymbolAddress returns for the current device (env_scratch_init).
// Every partial is written by k_env_latlong before k_env_finish reads it, so
// nothing in the array has to start at zero; calls that overlap on one GPU
// share it and race.
static __device__ __align__(256) unsigned char scratch[kEnvScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

hipError_t env_scratch_init(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return e;
  if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current
  void* p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!p) return hipErrorInvalidSymbol;
  g_scratch_ptr[device] = p;
  return hipSuccess;
}

void* env_scratch_ptr(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

}  // namespace shb
