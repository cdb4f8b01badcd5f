// Environment SH: projection of an equirectangular (LatLong) radiance map onto
// the nine real SH basis functions, src/scene.cpp:31-71 (ProjectLatLongMap).
//
// Per texel (x, y) of a w x h map: theta = (y + 0.5) pi / h from the zenith
// (+Y), phi = (x + 0.5) 2 pi / w from +Z towards +X, direction
// (sin t sin p, cos t, sin t cos p), weight d_omega = (2 pi / w)(pi / h) sin t;
// rows with sin t < 1e-4 are skipped, as in the reference. Every per-texel
// term (radiance * d_omega) * Y_i is the reference's float product; only the
// summation differs: FP32 fused sums over a few rows, then FP64 per thread,
// per block and across blocks in a fixed order (deterministic, and far
// closer to the exact sum than the reference's serial float sum).
//
// HBM-bound: each texel's channels are read once (4 * channels bytes); one
// thread owns one column of a block's row strip, so sin/cos of phi are
// computed once per thread and the row terms (sin t, cos t, d_omega) come
// from an LDS table filled once per block.
#include <algorithm>

#include "bake_device.h"
#include "kernels.h"

namespace shb {

namespace {

constexpr float kEnvPi = 3.14159265359f;     // scene.cpp:12
constexpr float kEnvTwoPi = 6.28318530718f;  // scene.cpp:13
constexpr int kEnvThreads = 256;
constexpr int kEnvMaxRows = 256;  // rows per block (LDS table size)
constexpr int kEnvFlushRows = 8;  // rows summed in FP32 before the FP64 add

__device__ __forceinline__ double block_sum(double v, double* red) {
  // wave64 shuffle tree, then the waves' sums in order (deterministic)
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < kEnvThreads / 64; ++k) s += red[k];
  return s;
}

}  // namespace

__global__ void __launch_bounds__(kEnvThreads)
    k_env_latlong(const float* __restrict__ px, int w, int h, int ch, int rows_per_block,
                  double* __restrict__ partials) {
  __shared__ float row_st[kEnvMaxRows], row_ct[kEnvMaxRows], row_dw[kEnvMaxRows];
  __shared__ double red[kEnvThreads / 64];
  const float d_phi = kEnvTwoPi / (float)w;
  const float d_theta = kEnvPi / (float)h;
  const int y0 = blockIdx.y * rows_per_block;
  const int nrows = min(rows_per_block, h - y0);
  for (int r = threadIdx.x; r < nrows; r += kEnvThreads) {
    const float theta = ((float)(y0 + r) + 0.5f) * d_theta;
    const float st = sinf(theta);
    row_st[r] = st;
    row_ct[r] = cosf(theta);
    row_dw[r] = st < 1e-4f ? 0.0f : d_phi * d_theta * st;  // 0 marks a skipped row
  }
  __syncthreads();
  double acc[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) acc[i] = 0.0;
  const int x = blockIdx.x * kEnvThreads + threadIdx.x;
  if (x < w) {
    const float phi = ((float)x + 0.5f) * d_phi;
    const float sp = sinf(phi), cp = cosf(phi);
    const float* p = px + ((size_t)y0 * w + x) * ch;
    const size_t row_stride = (size_t)w * ch;
    for (int r0 = 0; r0 < nrows; r0 += kEnvFlushRows) {
      float part[27];
#pragma unroll
      for (int i = 0; i < 27; ++i) part[i] = 0.0f;
      const int r1 = min(nrows, r0 + kEnvFlushRows);
      for (int r = r0; r < r1; ++r) {
        const float dw = row_dw[r];
        if (dw == 0.0f) continue;
        const float* q = p + r * row_stride;
        const float st = row_st[r];
        const f3 dir = {st * sp, row_ct[r], st * cp};
        const float cr = q[0] * dw, cg = q[1] * dw, cb = q[2] * dw;
        float Y[9];
        sh_basis(dir, Y);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
          part[3 * i] = fmaf(cr, Y[i], part[3 * i]);
          part[3 * i + 1] = fmaf(cg, Y[i], part[3 * i + 1]);
          part[3 * i + 2] = fmaf(cb, Y[i], part[3 * i + 2]);
        }
      }
#pragma unroll
      for (int i = 0; i < 27; ++i) acc[i] += (double)part[i];
    }
  }
  double* out = partials + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 27;
  for (int i = 0; i < 27; ++i) {
    const double s = block_sum(acc[i], red);
    if (threadIdx.x == 0) out[i] = s;
  }
}

// Sums the per-block partials in block order: out[i] = sum_b partials[b][i].
__global__ void __launch_bounds__(kEnvThreads)
    k_env_finish(const double* __restrict__ partials, int nblocks, double* __restrict__ out) {
  __shared__ double red[kEnvThreads / 64];
  for (int i = 0; i < 27; ++i) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kEnvThreads) s += partials[(size_t)b * 27 + i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[i] = s;
  }
}

EnvGrid env_grid(int w, int h) {
  EnvGrid g;
  g.gx = (w + kEnvThreads - 1) / kEnvThreads;
  // about 4096 blocks (16 per CU), each a strip of at most kEnvMaxRows rows
  long long want = ((long long)h * g.gx + 4095) / 4096;
  g.rows_per_block = (int)std::min<long long>(kEnvMaxRows, std::max<long long>(1, want));
  g.gy = (h + g.rows_per_block - 1) / g.rows_per_block;
  return g;
}

hipError_t launch_env_latlong(const float* px, int w, int h, int ch, const EnvGrid& g,
                              double* partials, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_env_latlong, dim3(g.gx, g.gy), dim3(kEnvThreads), 0, st, px, w, h, ch,
                     g.rows_per_block, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_env_finish, dim3(1), dim3(kEnvThreads), 0, st, partials, g.gx * g.gy,
                     out);
  return hipGetLastError();
}

}  // namespace shb
