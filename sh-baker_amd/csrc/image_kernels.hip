// Lightmap-space kernels: UV rasterizer, validity mask, box downsample,
// dilation, valid-texel compaction, and the exclusive scans they use (a
// single-block scan of per-block partial sums: no cross-block waits). Compiled with -ffp-contract=off; every
// float expression keeps the reference's operation order so results are bit
// exact against a serial restatement.
#include "kernels.h"

namespace shb {

// ------------------------------------------------------------- rasterizer
// src/rasterizer.cpp:101-112: texel bbox of a triangle's lightmap uvs.
__device__ __forceinline__ bool tri_bbox(const RasterArgs& a, uint32_t t, int& x0, int& x1,
                                         int& y0, int& y1, float2& A, float2& B, float2& C) {
  uint32_t i0 = a.tri_vidx[3 * t], i1 = a.tri_vidx[3 * t + 1], i2 = a.tri_vidx[3 * t + 2];
  A = make_float2(a.luv[2 * i0], a.luv[2 * i0 + 1]);
  B = make_float2(a.luv[2 * i1], a.luv[2 * i1 + 1]);
  C = make_float2(a.luv[2 * i2], a.luv[2 * i2 + 1]);
  float min_u = fminf(fminf(A.x, B.x), C.x), max_u = fmaxf(fmaxf(A.x, B.x), C.x);
  float min_v = fminf(fminf(A.y, B.y), C.y), max_v = fmaxf(fmaxf(A.y, B.y), C.y);
  x0 = max(0, (int)(min_u * (float)a.W));
  x1 = min(a.W - 1, (int)ceilf(max_u * (float)a.W));
  y0 = max(0, (int)(min_v * (float)a.H));
  y1 = min(a.H - 1, (int)ceilf(max_v * (float)a.H));
  return x1 >= x0 && y1 >= y0;
}

// src/rasterizer.cpp:18-33 Barycentric.
__device__ __forceinline__ bool barycentric(float2 p, float2 a, float2 b, float2 c, float& u,
                                            float& v, float& w) {
  float2 v0 = make_float2(b.x - a.x, b.y - a.y);
  float2 v1 = make_float2(c.x - a.x, c.y - a.y);
  float2 v2 = make_float2(p.x - a.x, p.y - a.y);
  float d00 = v0.x * v0.x + v0.y * v0.y;
  float d01 = v0.x * v1.x + v0.y * v1.y;
  float d11 = v1.x * v1.x + v1.y * v1.y;
  float d20 = v2.x * v0.x + v2.y * v0.y;
  float d21 = v2.x * v1.x + v2.y * v1.y;
  float denom = d00 * d11 - d01 * d01;
  if (fabsf(denom) < 1e-8f) return false;
  v = (d11 * d20 - d01 * d21) / denom;
  w = (d00 * d21 - d01 * d20) / denom;
  u = 1.0f - v - w;
  return (v >= 0.0f) && (w >= 0.0f) && (u >= 0.0f);
}

__global__ void k_raster_count(RasterArgs a, uint32_t* counts) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.num_tris) return;
  int x0, x1, y0, y1;
  float2 A, B, C;
  uint32_t c = 0;
  if (tri_bbox(a, t, x0, x1, y0, y1, A, B, C))
    c = (uint32_t)(x1 - x0 + 1) * (uint32_t)(y1 - y0 + 1);
  counts[t] = c;
}

// One thread per (triangle, bbox texel): coverage test, last regular writer
// by atomicMax of the serial triangle index (geometry-major order).
__global__ void k_raster_cover(RasterArgs a, const unsigned long long* scan,
                               unsigned long long total) {
  unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  for (unsigned long long item = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
       item < total; item += stride) {
    uint32_t lo = 0, hi = a.num_tris;  // find t: scan[t] <= item < scan[t+1]
    while (hi - lo > 1) {
      uint32_t mid = (lo + hi) >> 1;
      if (scan[mid] <= item)
        lo = mid;
      else
        hi = mid;
    }
    uint32_t t = lo;
    int x0, x1, y0, y1;
    float2 A, B, C;
    tri_bbox(a, t, x0, x1, y0, y1, A, B, C);
    uint32_t local = (uint32_t)(item - scan[t]);
    uint32_t bw = (uint32_t)(x1 - x0 + 1);
    int x = x0 + (int)(local % bw), y = y0 + (int)(local / bw);
    float2 p = make_float2(((float)x + 0.5f) / (float)a.W, ((float)y + 0.5f) / (float)a.H);
    float u, v, w;
    if (!barycentric(p, A, B, C, u, v, w)) continue;
    a.tri_hit[t] = 1u;
    atomicMax(a.win + (size_t)y * a.W + x, t + 1u);
  }
}

// src/rasterizer.cpp:160-205: a triangle that covered no texel centre
// writes its centroid texel if that texel is still empty; with serial
// order that is the first such triangle when no regular writer exists.
__global__ void k_raster_fallback(RasterArgs a) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.num_tris || a.tri_hit[t]) return;
  uint32_t i0 = a.tri_vidx[3 * t], i1 = a.tri_vidx[3 * t + 1], i2 = a.tri_vidx[3 * t + 2];
  float cu = (a.luv[2 * i0] + a.luv[2 * i1] + a.luv[2 * i2]) / 3.0f;
  float cv = (a.luv[2 * i0 + 1] + a.luv[2 * i1 + 1] + a.luv[2 * i2 + 1]) / 3.0f;
  int cx = (int)(cu * (float)a.W), cy = (int)(cv * (float)a.H);
  if (cx >= 0 && cx < a.W && cy >= 0 && cy < a.H) atomicMin(a.fb + (size_t)cy * a.W + cx, t);
}

__device__ __forceinline__ f3 ld3(const float* p, uint32_t i) {
  return {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
}

// src/rasterizer.cpp:133-151: interpolated world position / normal /
// Gram-Schmidt tangent, written as the 48-byte surface point.
__global__ void k_raster_resolve(RasterArgs a) {
  size_t n = (size_t)a.W * a.H;
  size_t px = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (px >= n) return;
  uint32_t wv = a.win[px];
  uint32_t fb = a.fb[px];
  float4* out = a.surface + 3 * px;
  uint32_t t;
  float u, v, w;
  if (wv != 0u) {
    t = wv - 1u;
    int x = (int)(px % (size_t)a.W), y = (int)(px / (size_t)a.W);
    uint32_t i0 = a.tri_vidx[3 * t], i1 = a.tri_vidx[3 * t + 1], i2 = a.tri_vidx[3 * t + 2];
    float2 A = make_float2(a.luv[2 * i0], a.luv[2 * i0 + 1]);
    float2 B = make_float2(a.luv[2 * i1], a.luv[2 * i1 + 1]);
    float2 C = make_float2(a.luv[2 * i2], a.luv[2 * i2 + 1]);
    float2 p = make_float2(((float)x + 0.5f) / (float)a.W, ((float)y + 0.5f) / (float)a.H);
    barycentric(p, A, B, C, u, v, w);
  } else if (fb != 0xffffffffu) {
    t = fb;
    u = v = w = 1.0f / 3.0f;
  } else {
    out[0] = make_float4(0.f, 0.f, 0.f, 0.f);
    out[1] = make_float4(0.f, 0.f, 0.f, 0.f);
    out[2] = make_float4(0.f, 0.f, as_float(0xffffffffu), 0.f);
    return;
  }
  uint32_t i0 = a.tri_vidx[3 * t], i1 = a.tri_vidx[3 * t + 1], i2 = a.tri_vidx[3 * t + 2];
  f3 pos = ld3(a.wpos, i0) * u + ld3(a.wpos, i1) * v + ld3(a.wpos, i2) * w;
  f3 nrm = normalized(ld3(a.wnrm, i0) * u + ld3(a.wnrm, i1) * v + ld3(a.wnrm, i2) * w);
  const float* T0 = a.wtan + 4 * (size_t)i0;
  const float* T1 = a.wtan + 4 * (size_t)i1;
  const float* T2 = a.wtan + 4 * (size_t)i2;
  float tx = T0[0] * u + T1[0] * v + T2[0] * w;
  float ty = T0[1] * u + T1[1] * v + T2[1] * w;
  float tz = T0[2] * u + T1[2] * v + T2[2] * w;
  float tw = T0[3] * u + T1[3] * v + T2[3] * w;
  f3 t3 = {tx, ty, tz};
  t3 = normalized(t3 - nrm * dot(nrm, t3));
  out[0] = make_float4(pos.x, pos.y, pos.z, nrm.x);
  out[1] = make_float4(nrm.y, nrm.z, t3.x, t3.y);
  out[2] = make_float4(t3.z, tw, as_float((uint32_t)a.tri_mat[t]), 0.f);
}

// Exclusive scan of n uint32 counts into n + 1 uint64 offsets (out[n] =
// total) by one block of 1024 threads: each thread sums a contiguous chunk,
// the 1024 chunk sums are scanned in LDS, then every chunk is rewritten.
constexpr int kScanThreads = 1024;
__global__ void __launch_bounds__(kScanThreads) k_scan_u32_to_u64(const uint32_t* in, uint32_t n,
                                                                 unsigned long long* out) {
  __shared__ unsigned long long part[kScanThreads];
  const uint32_t tid = threadIdx.x;
  const uint32_t chunk = (n + kScanThreads - 1) / kScanThreads;
  const uint32_t b = tid * chunk, e = min(n, b + chunk);
  unsigned long long acc = 0;
  for (uint32_t k = b; k < e; ++k) acc += in[k];
  part[tid] = acc;
  __syncthreads();
  if (tid == 0) {
    unsigned long long run = 0;
    for (int k = 0; k < kScanThreads; ++k) {
      unsigned long long v = part[k];
      part[k] = run;
      run += v;
    }
    out[n] = run;
  }
  __syncthreads();
  unsigned long long run = part[tid];
  for (uint32_t k = b; k < e; ++k) {
    out[k] = run;
    run += in[k];
  }
}

hipError_t launch_raster(const RasterArgs& a, uint32_t* counts, unsigned long long* scan,
                         hipStream_t st) {
  const uint32_t n = a.num_tris;
  size_t npx = (size_t)a.W * a.H;
  hipError_t e;
  if ((e = launch_fill_u32(a.win, 0u, npx, st)) != hipSuccess) return e;
  if ((e = launch_fill_u32(a.fb, 0xffffffffu, npx, st)) != hipSuccess) return e;
  if (n > 0) {
    if ((e = launch_fill_u32(a.tri_hit, 0u, n, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_raster_count, dim3((n + 255) / 256), dim3(256), 0, st, a, counts);
    hipLaunchKernelGGL(k_scan_u32_to_u64, dim3(1), dim3(kScanThreads), 0, st, counts, n, scan);
    unsigned long long total = 0;
    if ((e = hipMemcpyAsync(&total, scan + n, sizeof(total), hipMemcpyDeviceToHost, st)) != hipSuccess)
      return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    if (total > 0) {
      unsigned long long blocks = (total + 255) / 256;
      int grid = (int)(blocks < 65536ull ? blocks : 65536ull);
      hipLaunchKernelGGL(k_raster_cover, dim3(grid), dim3(256), 0, st, a, scan, total);
    }
    hipLaunchKernelGGL(k_raster_fallback, dim3((n + 255) / 256), dim3(256), 0, st, a);
  }
  if (npx > 0)
    hipLaunchKernelGGL(k_raster_resolve, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

// ------------------------------------------------------------------ masks
// src/rasterizer.cpp:213-220 CreateValidityMask.
__global__ void k_validity(const float4* surface, size_t n, uint8_t* mask) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  mask[i] = (int)as_uint(surface[3 * i + 2].z) >= 0 ? 1 : 0;
}

hipError_t launch_validity_mask(const float4* surface, size_t n, uint8_t* mask, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_validity, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, surface, n,
                     mask);
  return hipGetLastError();
}

// src/rasterizer.cpp:222-241 DownsampleValidityMask (w, h: output size).
__global__ void k_downsample_mask(const uint8_t* in, int w, int h, int scale, uint8_t* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w * h) return;
  int x = i % w, y = i / w;
  int stride = w * scale;
  uint8_t any = 0;
  for (int sy = y * scale; sy < (y + 1) * scale; ++sy)
    for (int sx = x * scale; sx < (x + 1) * scale; ++sx) any |= in[(size_t)sy * stride + sx] ? 1 : 0;
  out[i] = any;
}

hipError_t launch_downsample_mask(const uint8_t* in, int w, int h, int scale, uint8_t* out,
                                  hipStream_t st) {
  if (w * h == 0) return hipSuccess;
  hipLaunchKernelGGL(k_downsample_mask, dim3((w * h + 255) / 256), dim3(256), 0, st, in, w, h,
                     scale, out);
  return hipGetLastError();
}

// src/baker.cpp:151-182 DownsampleTexture: ss x ss box sum (dy-major,
// dx-minor) times 1/count, invalid texels included. One thread per output
// channel; w, h: output size.
__global__ void k_downsample(const float* in, int w, int h, int scale, int ch, float* out) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  size_t n = (size_t)w * h * ch;
  if (i >= n) return;
  int c = (int)(i % ch);
  size_t px = i / ch;
  int x = (int)(px % w), y = (int)(px / w);
  int in_w = w * scale;
  float acc = 0.f;
  int count = 0;
  for (int dy = 0; dy < scale; ++dy)
    for (int dx = 0; dx < scale; ++dx) {
      int sx = x * scale + dx, sy = y * scale + dy;
      acc += in[((size_t)sy * in_w + sx) * ch + c];
      count++;
    }
  out[i] = acc * (1.f / float(count));
}

hipError_t launch_downsample(const float* in, int w, int h, int scale, int ch, float* out,
                             hipStream_t st) {
  size_t n = (size_t)w * h * ch;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_downsample, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, w, h,
                     scale, ch, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------- dilation
// src/dilation.cpp:7-73, one Jacobi pass: an invalid texel with valid
// 8-neighbours becomes their average (neighbours summed dy-major, dx-minor,
// then times 1/count) and valid in the next mask; all other texels copy.
// 16x16 tiles: the mask tile with its one-texel halo is staged in LDS once,
// then each channel's tile + halo is staged and reduced.
constexpr int kTile = 16;
__global__ void __launch_bounds__(256) k_dilate_pass(const float* src, float* dst,
                                                     const uint8_t* msrc, uint8_t* mdst,
                                                     int w, int h, int ch) {
  __shared__ uint8_t m[kTile + 2][kTile + 2];
  __shared__ float v[kTile + 2][kTile + 2];
  const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
  const int x0 = blockIdx.x * kTile, y0 = blockIdx.y * kTile;
  const int x = x0 + tx, y = y0 + ty;
  for (int k = threadIdx.x; k < (kTile + 2) * (kTile + 2); k += blockDim.x) {
    int lx = k % (kTile + 2), ly = k / (kTile + 2);
    int gx = x0 + lx - 1, gy = y0 + ly - 1;
    m[ly][lx] = (gx >= 0 && gx < w && gy >= 0 && gy < h) ? msrc[(size_t)gy * w + gx] : 0;
  }
  __syncthreads();
  const bool inside = x < w && y < h;
  const bool self = inside && m[ty + 1][tx + 1];
  int count = 0;
  if (inside && !self) {
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        if (dx == 0 && dy == 0) continue;
        count += m[ty + 1 + dy][tx + 1 + dx] ? 1 : 0;
      }
  }
  for (int c = 0; c < ch; ++c) {
    __syncthreads();
    for (int k = threadIdx.x; k < (kTile + 2) * (kTile + 2); k += blockDim.x) {
      int lx = k % (kTile + 2), ly = k / (kTile + 2);
      int gx = x0 + lx - 1, gy = y0 + ly - 1;
      v[ly][lx] = (gx >= 0 && gx < w && gy >= 0 && gy < h) ? src[((size_t)gy * w + gx) * ch + c] : 0.f;
    }
    __syncthreads();
    if (!inside) continue;
    float out = v[ty + 1][tx + 1];
    if (count > 0) {
      float acc = 0.f;
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          if (dx == 0 && dy == 0) continue;
          if (m[ty + 1 + dy][tx + 1 + dx]) acc += v[ty + 1 + dy][tx + 1 + dx];
        }
      out = acc * (1.0f / count);
    }
    dst[((size_t)y * w + x) * ch + c] = out;
  }
  if (inside) mdst[(size_t)y * w + x] = (self || count > 0) ? 1 : 0;
}

// The same pass with every channel of the tile and its halo staged in LDS
// at once: rows of the interleaved image are contiguous (18 texels x ch
// floats), so loads and stores are coalesced and each texel is read from
// memory once per pass. Same arithmetic and order as k_dilate_pass.
constexpr int kDilateMaxCh = 32;  // LDS: 18 x 18 x ch floats (40.5 KB at 32)
__global__ void __launch_bounds__(256) k_dilate_tile(const float* src, float* dst,
                                                     const uint8_t* msrc, uint8_t* mdst, int w,
                                                     int h, int ch) {
  extern __shared__ float tile[];  // (kTile + 2)^2 texels x ch
  __shared__ uint8_t m[kTile + 2][kTile + 2];
  constexpr int E = kTile + 2;
  const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
  const int x0 = blockIdx.x * kTile, y0 = blockIdx.y * kTile;
  const int x = x0 + tx, y = y0 + ty;
  for (int k = threadIdx.x; k < E * E; k += blockDim.x) {
    int lx = k % E, ly = k / E;
    int gx = x0 + lx - 1, gy = y0 + ly - 1;
    m[ly][lx] = (gx >= 0 && gx < w && gy >= 0 && gy < h) ? msrc[(size_t)gy * w + gx] : 0;
  }
  // rows of the tile + halo: E texels x ch floats each, contiguous in src
  const int rowf = E * ch;
  for (int k = threadIdx.x; k < E * rowf; k += blockDim.x) {
    const int ly = k / rowf, r = k - ly * rowf;
    const int lx = r / ch;
    const int gx = x0 + lx - 1, gy = y0 + ly - 1;
    tile[k] = (gx >= 0 && gx < w && gy >= 0 && gy < h)
                  ? src[((size_t)gy * w + (size_t)(x0 - 1)) * ch + r]
                  : 0.f;
  }
  __syncthreads();
  const bool inside = x < w && y < h;
  const bool self = inside && m[ty + 1][tx + 1];
  int count = 0;
  if (inside && !self) {
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        if (dx == 0 && dy == 0) continue;
        count += m[ty + 1 + dy][tx + 1 + dx] ? 1 : 0;
      }
  }
  float out[kDilateMaxCh];
  const float* cen = tile + ((ty + 1) * E + (tx + 1)) * ch;
  for (int c = 0; c < ch; ++c) {
    float o = cen[c];
    if (count > 0) {
      float acc = 0.f;
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          if (dx == 0 && dy == 0) continue;
          if (m[ty + 1 + dy][tx + 1 + dx]) acc += tile[((ty + 1 + dy) * E + (tx + 1 + dx)) * ch + c];
        }
      o = acc * (1.0f / count);
    }
    out[c] = o;
  }
  __syncthreads();  // every thread has read its neighbours
  float* own = tile + ((ty + 1) * E + (tx + 1)) * ch;
  for (int c = 0; c < ch; ++c) own[c] = out[c];
  __syncthreads();
  // interior rows back to dst, contiguous kTile texels x ch floats
  const int rin = kTile * ch;
  for (int k = threadIdx.x; k < kTile * rin; k += blockDim.x) {
    const int ly = k / rin, r = k - ly * rin;
    const int lx = r / ch;
    const int gx = x0 + lx, gy = y0 + ly;
    if (gx < w && gy < h)
      dst[((size_t)gy * w + (size_t)x0) * ch + r] = tile[((ly + 1) * E + 1) * ch + r];
  }
  if (inside) mdst[(size_t)y * w + x] = (self || count > 0) ? 1 : 0;
}

// Ping-pongs between (a, ma) and (b, mb); *result_in_a tells where the final
// pass landed. Passes after the first unchanged pass would copy only, so
// running them all gives the reference's early-exit result.
hipError_t launch_dilate(float* a, float* b, uint8_t* ma, uint8_t* mb, int w, int h, int ch,
                         int passes, uint32_t* changed, bool* result_in_a, hipStream_t st) {
  (void)changed;
  bool in_a = true;
  dim3 grid((w + kTile - 1) / kTile, (h + kTile - 1) / kTile);
  const bool tiled = ch >= 1 && ch <= kDilateMaxCh;
  const size_t lds = tiled ? (size_t)(kTile + 2) * (kTile + 2) * ch * sizeof(float) : 0;
  for (int p = 0; p < passes; ++p) {
    float* s0 = in_a ? a : b;
    float* d0 = in_a ? b : a;
    uint8_t* ms = in_a ? ma : mb;
    uint8_t* md = in_a ? mb : ma;
    if (tiled)
      hipLaunchKernelGGL(k_dilate_tile, grid, dim3(kTile * kTile), lds, st, s0, d0, ms, md, w, h,
                         ch);
    else
      hipLaunchKernelGGL(k_dilate_pass, grid, dim3(kTile * kTile), 0, st, s0, d0, ms, md, w, h, ch);
    in_a = !in_a;
  }
  *result_in_a = in_a;
  return hipGetLastError();
}

// ------------------------------------------------------------- compaction
// Ascending list of the set mask bytes: per-block counts, a single-block scan
// of the counts, then an ordered in-block scatter (wave ballots + LDS).
constexpr int kCompactBlock = 256;

uint32_t compact_blocks(uint32_t n) { return (n + kCompactBlock - 1) / kCompactBlock; }

__device__ __forceinline__ uint32_t block_exclusive_flags(bool flag, uint32_t* total,
                                                          uint32_t* wave_sums) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t m = __ballot(flag);
  uint32_t in_wave = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) wave_sums[wave] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (int w = 0; w < kCompactBlock / 64; ++w) {
    if (w < wave) before += wave_sums[w];
    all += wave_sums[w];
  }
  *total = all;
  return before + in_wave;
}

__global__ void __launch_bounds__(kCompactBlock) k_compact_count(const uint8_t* mask, uint32_t n,
                                                                 uint32_t* block_counts) {
  __shared__ uint32_t wave_sums[kCompactBlock / 64];
  uint32_t i = blockIdx.x * kCompactBlock + threadIdx.x;
  uint32_t total;
  block_exclusive_flags(i < n && mask[i] != 0, &total, wave_sums);
  if (threadIdx.x == 0) block_counts[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_u32_inplace(uint32_t* v, uint32_t n,
                                                                   uint32_t* total) {
  __shared__ uint32_t part[kScanThreads];
  const uint32_t tid = threadIdx.x;
  const uint32_t chunk = (n + kScanThreads - 1) / kScanThreads;
  const uint32_t b = tid * chunk, e = min(n, b + chunk);
  uint32_t acc = 0;
  for (uint32_t k = b; k < e; ++k) acc += v[k];
  part[tid] = acc;
  __syncthreads();
  if (tid == 0) {
    uint32_t run = 0;
    for (int k = 0; k < kScanThreads; ++k) {
      uint32_t x = part[k];
      part[k] = run;
      run += x;
    }
    *total = run;
  }
  __syncthreads();
  uint32_t run = part[tid];
  for (uint32_t k = b; k < e; ++k) {
    uint32_t x = v[k];
    v[k] = run;
    run += x;
  }
}

__global__ void __launch_bounds__(kCompactBlock) k_compact_scatter(const uint8_t* mask, uint32_t n,
                                                                   const uint32_t* block_off,
                                                                   uint32_t* list) {
  __shared__ uint32_t wave_sums[kCompactBlock / 64];
  uint32_t i = blockIdx.x * kCompactBlock + threadIdx.x;
  bool f = i < n && mask[i] != 0;
  uint32_t total;
  uint32_t at = block_exclusive_flags(f, &total, wave_sums);
  if (f) list[block_off[blockIdx.x] + at] = i;
}

hipError_t launch_compact_valid(const uint8_t* mask, uint32_t n, uint32_t* list,
                                uint32_t* block_tmp, hipStream_t st) {
  uint32_t nb = compact_blocks(n);
  if (nb == 0) return launch_fill_u32(list + n, 0u, 1, st);
  hipLaunchKernelGGL(k_compact_count, dim3(nb), dim3(kCompactBlock), 0, st, mask, n, block_tmp);
  hipLaunchKernelGGL(k_scan_u32_inplace, dim3(1), dim3(kScanThreads), 0, st, block_tmp, nb,
                     list + n);
  hipLaunchKernelGGL(k_compact_scatter, dim3(nb), dim3(kCompactBlock), 0, st, mask, n, block_tmp,
                     list);
  return hipGetLastError();
}

}  // namespace shb
