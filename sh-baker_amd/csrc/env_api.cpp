// Environment SH entry points (src/scene.cpp:22-160, ProjectEnvironmentToSH):
// the LatLong projection runs on the device (env_kernels.hip); the Preetham
// sky is a fixed 64 x 128 quadrature evaluated on the host.
#include <cmath>
#include <cstring>
#include <string>

#include "api_util.h"
#include "kernels.h"

using namespace shb;

namespace {

// scene.cpp:15-20: Perez coefficients of a clear sky (turbidity ~2.5)
constexpr float kPerezA = -1.0f, kPerezB = -0.32f, kPerezC = 10.0f, kPerezD = -3.0f,
                kPerezE = 0.45f;
constexpr float kEnvPi = 3.14159265359f, kEnvTwoPi = 6.28318530718f;

struct V3 {
  float x, y, z;
};

// scene.cpp:75-94 EvaluatePerez (view zenith clamped at 0.001, sun angle by acos)
float perez(V3 d, V3 s) {
  const float cos_zenith = std::max(0.001f, d.y);
  const float cos_sun = d.x * s.x + (d.y * s.y + d.z * s.z);
  const float gamma = std::acos(std::min(1.0f, std::max(-1.0f, cos_sun)));
  const float f_zenith = 1.0f + kPerezA * std::exp(kPerezB / cos_zenith);
  const float f_sun = 1.0f + kPerezC * std::exp(kPerezD * gamma) + kPerezE * cos_sun * cos_sun;
  return f_zenith * f_sun;
}

// src/sh_coeffs.cpp:50-78 AccumulateRadiance: sh[3 i + c] += rad[c] * Y_i(d)
void accumulate(const float rad[3], V3 d, float* sh) {
  float Y[9];
  Y[0] = 0.282095f;
  Y[1] = 0.488603f * d.y;
  Y[2] = 0.488603f * d.z;
  Y[3] = 0.488603f * d.x;
  Y[4] = 1.092548f * d.x * d.y;
  Y[5] = 1.092548f * d.y * d.z;
  Y[6] = 0.315392f * (3.0f * d.z * d.z - 1.0f);
  Y[7] = 1.092548f * d.x * d.z;
  Y[8] = 0.546274f * (d.x * d.x - d.y * d.y);
  for (int i = 0; i < 9; ++i)
    for (int c = 0; c < 3; ++c) sh[3 * i + c] += rad[c] * Y[i];
}

}  // namespace

extern "C" {

int shb_env_sh_latlong(shb_ctx* h, int32_t width, int32_t height, int32_t channels,
                       const float* pixels, float* sh_out, double* seconds_device) {
  if (!h) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  if (!sh_out) return fail(c, SHB_ERR_INVALID_ARG, "sh_out is NULL");
  if (width <= 0 || height <= 0 || channels < 3)
    return fail(c, SHB_ERR_INVALID_ARG, "bad environment map size or channel count (< 3)");
  if ((size_t)width * height > (size_t)1 << 31)
    return fail(c, SHB_ERR_LIMIT, "environment map above 2^31 texels");
  if (!pixels && (width != c.env_w || height != c.env_h || channels != c.env_ch))
    return fail(c, SHB_ERR_INVALID_ARG,
                "pixels is NULL but no map of this size is resident from a previous call");
  SHB_TRY(c, hipSetDevice(c.device));
  const size_t bytes = (size_t)width * height * channels * sizeof(float);
  if (pixels) {
    SHB_TRY(c, c.d_env.reserve(bytes));
    SHB_TRY(c, hipMemcpyAsync(c.d_env.ptr, pixels, bytes, hipMemcpyHostToDevice, c.stream));
    c.env_w = width;
    c.env_h = height;
    c.env_ch = channels;
  }
  const EnvGrid g = env_grid(width, height);
  SHB_TRY(c, c.d_env_part.reserve(((size_t)g.gx * g.gy + 1) * 27 * sizeof(double)));
  double* part = c.d_env_part.as<double>();
  double* out = part + (size_t)g.gx * g.gy * 27;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  SHB_TRY(c, hipEventCreate(&e0));
  hipError_t err = hipEventCreate(&e1);
  if (err == hipSuccess) err = hipEventRecord(e0, c.stream);
  if (err == hipSuccess)
    err = launch_env_latlong(c.d_env.as<float>(), width, height, channels, g, part, out,
                             nullptr, c.stream);
  if (err == hipSuccess) err = hipEventRecord(e1, c.stream);
  double sums[27];
  if (err == hipSuccess)
    err = hipMemcpyAsync(sums, out, sizeof(sums), hipMemcpyDeviceToHost, c.stream);
  if (err == hipSuccess) err = hipStreamSynchronize(c.stream);
  float ms = 0.f;
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  SHB_TRY(c, err);
  for (int i = 0; i < 27; ++i) sh_out[i] = (float)sums[i];
  if (seconds_device) *seconds_device = ms * 1e-3;
  return SHB_OK;
}

int shb_env_sh_latlong_noalloc(shb_ctx* h, int32_t width, int32_t height, int32_t channels,
                               const float* pixels_device, float* sh_out_device) {
  if (!h) return SHB_ERR_INVALID_ARG;
  Ctx& c = h->c;
  if (!pixels_device || !sh_out_device)
    return fail(c, SHB_ERR_INVALID_ARG, "pixels_device and sh_out_device must not be NULL");
  if (width <= 0 || height <= 0 || channels < 3)
    return fail(c, SHB_ERR_INVALID_ARG, "bad environment map size or channel count (< 3)");
  if ((size_t)width * height > (size_t)1 << 31)
    return fail(c, SHB_ERR_LIMIT, "environment map above 2^31 texels");
  const EnvGrid g = env_grid(width, height);
  const size_t nblocks = (size_t)g.gx * g.gy;
  if ((nblocks + 1) * 27 * sizeof(double) > kEnvScratchBytes)
    return fail(c, SHB_ERR_LIMIT,
                "environment map needs " + std::to_string(nblocks) +
                    " partial-sum blocks; the fixed scratch holds " +
                    std::to_string(kEnvScratchBytes / (27 * sizeof(double)) - 1) +
                    " (use shb_env_sh_latlong)");
  SHB_TRY(c, hipSetDevice(c.device));
  // both buffers must be device memory of this GPU (a host pointer would fault)
  for (const void* p : {(const void*)pixels_device, (const void*)sh_out_device}) {
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess || at.type != hipMemoryTypeDevice ||
        at.device != c.device) {
      (void)hipGetLastError();  // clear the lookup's error state
      return fail(c, SHB_ERR_INVALID_ARG,
                  "pixels_device and sh_out_device must be device memory of the context's GPU");
    }
  }
  double* part = static_cast<double*>(env_scratch_ptr(c.device));
  if (!part) return fail(c, SHB_ERR_DEVICE, "the fixed scratch of this GPU was not set up");
  SHB_TRY(c, launch_env_latlong(pixels_device, width, height, channels, g, part, nullptr,
                                sh_out_device, c.stream));
  SHB_TRY(c, hipStreamSynchronize(c.stream));
  return SHB_OK;
}

int shb_env_sh_preetham(const float sun_direction[3], float* sh_out) {
  if (!sun_direction || !sh_out) return SHB_ERR_INVALID_ARG;
  // scene.cpp:96-150 ProjectPreethamToSH; the sun direction is normalised
  // as Eigen's normalized() does (a zero vector stays zero)
  const float* s = sun_direction;
  if (!std::isfinite(s[0]) || !std::isfinite(s[1]) || !std::isfinite(s[2]))
    return SHB_ERR_INVALID_ARG;
  const float n2 = s[0] * s[0] + (s[1] * s[1] + s[2] * s[2]);
  V3 sun = {s[0], s[1], s[2]};
  if (n2 > 0.0f) {
    const float n = std::sqrt(n2);
    sun = {s[0] / n, s[1] / n, s[2] / n};
  }
  const int steps_theta = 64, steps_phi = 128;
  const float d_theta = kEnvPi / steps_theta, d_phi = kEnvTwoPi / steps_phi;
  float zenith = perez({0.f, 1.f, 0.f}, sun);
  if (zenith <= 1e-6f) zenith = 1.0f;
  const float cap = zenith * 30.0f;
  std::memset(sh_out, 0, 27 * sizeof(float));
  for (int t = 0; t < steps_theta; ++t) {
    const float theta = (t + 0.5f) * d_theta;
    const float st = std::sin(theta);
    for (int p = 0; p < steps_phi; ++p) {
      const float phi = (p + 0.5f) * d_phi;
      const V3 d = {st * std::sin(phi), std::cos(theta), st * std::cos(phi)};
      if (d.y < 0.0f) continue;  // ground is black
      const float lum = std::min(perez(d, sun), cap);
      const float k = lum / zenith;
      const float dw = st * d_theta * d_phi;
      const float rad[3] = {0.2f * k * dw, 0.5f * k * dw, 0.9f * k * dw};
      accumulate(rad, d, sh_out);
    }
  }
  return SHB_OK;
}

}  // extern "C"
