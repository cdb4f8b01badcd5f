// shb_baker.hpp — C++ mirror of the reference's Baker surface over the C ABI
// of libshb_hip.so (include/shb.h). Header only, C++17, no Eigen.
//
// Same names, argument meaning and error behaviour as DaviesX/sh-baker's
// src/baker.h, src/rasterizer.h, src/dilation.h, src/occlusion.h and the
// types of src/scene.h / src/saver.h / src/sh_coeffs.h, so `main`-style
// callers and tests written against the reference compile against this
// header with only the namespace changed (sh_baker -> shb_baker). The vector
// types are small PODs with Eigen-style x()/y()/z()/w() accessors.
//
// Error behaviour (as the reference): no exceptions; a failed call logs to
// stderr and returns an empty result (FindOcclusion: std::nullopt).
//
// Each free function creates the GPU work on a per-thread context (device 0
// unless SetDefaultDevice was called first), uploads the scene, runs on the
// GPU and returns host results, like the reference, which builds its Embree
// scene inside BakeSHLightMap (src/baker.cpp:218-219).
#ifndef SHB_BAKER_HPP_
#define SHB_BAKER_HPP_

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <limits>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "shb.h"

namespace shb_baker {

// ------------------------------------------------------------ vectors
template <int N>
struct Vec {
  std::array<float, N> v{};
  Vec() = default;
  template <class... T, class = std::enable_if_t<sizeof...(T) == N>>
  Vec(T... a) : v{{static_cast<float>(a)...}} {}
  float& operator[](int i) { return v[i]; }
  float operator[](int i) const { return v[i]; }
  float x() const { return v[0]; }
  float y() const { return v[1]; }
  template <int M = N, class = std::enable_if_t<(M > 2)>>
  float z() const { return v[2]; }
  template <int M = N, class = std::enable_if_t<(M > 3)>>
  float w() const { return v[3]; }
  static Vec Zero() { return Vec(); }
  static Vec Ones() {
    Vec r;
    r.v.fill(1.f);
    return r;
  }
};
using Vector2f = Vec<2>;
using Vector3f = Vec<3>;
using Vector4f = Vec<4>;

// ------------------------------------------------------- scene (scene.h)
struct Texture {
  uint32_t width = 0;
  uint32_t height = 0;
  uint32_t channels = 0;
  std::vector<uint8_t> pixel_data;
  std::optional<std::filesystem::path> file_path;  // absolute, when loaded from a file
};

struct Texture32F {
  uint32_t width = 0;
  uint32_t height = 0;
  uint32_t channels = 0;
  std::vector<float> pixel_data;
  std::optional<std::filesystem::path> file_path;
};

struct Material {
  std::string name;
  Texture albedo;
  float emission_intensity = 0.0f;
};

// Row-major 3x4 object-to-world affine; identity by default.
struct Affine3f {
  std::array<float, 12> m{{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0}};
  static Affine3f Identity() { return Affine3f(); }
};

struct Geometry {
  std::vector<Vector3f> vertices;
  std::vector<Vector3f> normals;
  std::vector<Vector2f> texture_uvs;
  std::vector<Vector2f> lightmap_uvs;
  std::vector<Vector4f> tangents;  // xyz + w (sign)
  std::vector<uint32_t> indices;
  int material_id = -1;
  Affine3f transform = Affine3f::Identity();
};

struct Light {
  enum class Type { Point, Directional, Spot, Area };
  Type type = Type::Point;
  Vector3f position = Vector3f::Zero();
  Vector3f direction = Vector3f(0, 0, -1);
  Vector3f color = Vector3f::Ones();
  float intensity = 1.0f;
  float cos_inner_cone = 1.0f;
  float cos_outer_cone = 0.70710678118654752440f;  // cos(pi/4)
  // Area lights: the emitting geometry and its material, by index (the
  // reference keeps pointers plus geometry_index).
  int geometry_index = -1;
  int material_index = -1;
};

// ------------------------------------------------------- (sh_coeffs.h)
struct SHCoeffs {
  std::array<Vector3f, 9> coeffs{};  // L0, L1m1, L10, L11, L2m2, L2m1, L20, L21, L22
  SHCoeffs& operator+=(const SHCoeffs& o) {
    for (int i = 0; i < 9; ++i)
      for (int c = 0; c < 3; ++c) coeffs[i][c] += o.coeffs[i][c];
    return *this;
  }
};

// Replaces Environment (src/scene.h:90-107): an HDR LatLong texture or a
// Preetham sky from a sun direction, with its SH9 projection.
struct Environment {
  enum class Type { Texture, Preetham };
  Type type = Type::Preetham;
  Texture32F texture;
  float intensity_multiplier = 1.0f;
  Vector3f sun_direction = Vector3f(0, 1, 0);
  float sun_intensity = 1.0f;
  float turbidity = 2.5f;
  SHCoeffs sh_coeffs;
};

struct Scene {
  std::vector<Geometry> geometries;
  std::vector<Material> materials;
  std::vector<Light> lights;
  std::optional<Environment> environment;
};

// ---------------------------------------------- raster (rasterizer.h)
struct RasterConfig {
  int width = 1024;
  int height = 1024;
  int supersample_scale = 1;
};

struct SurfacePoint {
  Vector3f position = Vector3f::Zero();
  Vector3f normal = Vector3f::Zero();
  Vector4f tangent = Vector4f::Zero();
  int32_t material_id = -1;  // Invalid if < 0.
};

// ----------------------------------------------------------- (saver.h)
struct SHTexture {
  int width = 0;
  int height = 0;
  std::vector<SHCoeffs> pixels;
};

// ---------------------------------------------------- bake (baker.h)
struct BakeConfig {
  int samples = 128;          // Rays per texel
  int bounces = 3;            // Max path depth
  int num_light_samples = 1;  // Number of light samples for NEE (1..16)
  uint64_t seed = 12345;      // counter-based stream key (no reference field)
};

struct BakeResult {
  SHTexture sh_texture;
  Texture32F environment_visibility_texture;
};

// ------------------------------------------------- occlusion (occlusion.h)
struct Ray {
  Vector3f origin;
  Vector3f direction;
  float tnear = 0.0f;
  float tfar = std::numeric_limits<float>::infinity();
};

struct Occlusion {
  Vector3f position;
  Vector2f uv;
  Vector3f normal;
  int material_id;
};

// ==================================================================
// Implementation
// ==================================================================
namespace detail {

inline void log_error(const char* what, shb_ctx* c, int rc) {
  std::fprintf(stderr, "shb_baker: %s failed (%d): %s\n", what, rc,
               c ? shb_last_error(c) : "no GPU context");
}

// One GPU context (device + stream + buffers), like an RTCDevice + RTCScene.
class Context {
 public:
  explicit Context(int device) {
    int rc = shb_create(&ctx_, device);
    if (rc != SHB_OK) {
      log_error("shb_create", nullptr, rc);
      ctx_ = nullptr;
    }
  }
  ~Context() {
    if (ctx_) shb_destroy(ctx_);
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  shb_ctx* get() const { return ctx_; }

 private:
  shb_ctx* ctx_ = nullptr;
};

inline int& default_device() {
  static int d = 0;
  return d;
}

inline Context& thread_context() {
  thread_local std::unique_ptr<Context> c;
  if (!c) c.reset(new Context(default_device()));
  return *c;
}

// Flattens a Scene into the ABI's arrays (the vectors must outlive upload).
struct FlatScene {
  std::vector<float> pos, nrm, tuv, luv, tan;
  std::vector<uint32_t> idx;
  std::vector<shb_geometry> geos;
  std::vector<shb_material> mats;
  std::vector<uint8_t> pool;
  std::vector<shb_light> lights;
  shb_scene_desc desc{};

  explicit FlatScene(const Scene& s) {
    bool any_n = false, any_t = false, any_l = false, any_g = false;
    for (const Geometry& g : s.geometries) {
      any_n |= !g.normals.empty();
      any_t |= !g.texture_uvs.empty();
      any_l |= !g.lightmap_uvs.empty();
      any_g |= !g.tangents.empty();
    }
    for (const Geometry& g : s.geometries) {
      shb_geometry d{};
      d.vertex_offset = (uint32_t)(pos.size() / 3);
      d.vertex_count = (uint32_t)g.vertices.size();
      d.index_offset = (uint32_t)idx.size();
      d.index_count = (uint32_t)g.indices.size();
      d.material_id = g.material_id;
      const size_t nv = g.vertices.size();
      auto put = [nv](std::vector<float>& dst, bool any, const auto& src, int k, uint32_t flag,
                      uint32_t& flags) {
        if (!any) return;
        if (src.size() == nv) {
          for (const auto& e : src)
            for (int i = 0; i < k; ++i) dst.push_back(e[i]);
          flags |= flag;
        } else {
          dst.resize(dst.size() + nv * k, 0.f);
        }
      };
      for (const Vector3f& v : g.vertices)
        for (int i = 0; i < 3; ++i) pos.push_back(v[i]);
      put(nrm, any_n, g.normals, 3, SHB_GEOM_HAS_NORMALS, d.flags);
      put(tuv, any_t, g.texture_uvs, 2, SHB_GEOM_HAS_TEXTURE_UVS, d.flags);
      put(luv, any_l, g.lightmap_uvs, 2, SHB_GEOM_HAS_LIGHTMAP_UVS, d.flags);
      put(tan, any_g, g.tangents, 4, SHB_GEOM_HAS_TANGENTS, d.flags);
      idx.insert(idx.end(), g.indices.begin(), g.indices.end());
      std::memcpy(d.transform, g.transform.m.data(), sizeof(d.transform));
      geos.push_back(d);
    }
    for (const Material& m : s.materials) {
      shb_material d{};
      d.albedo_width = m.albedo.width;
      d.albedo_height = m.albedo.height;
      d.albedo_channels = m.albedo.pixel_data.empty() ? 0 : m.albedo.channels;
      d.albedo_offset = pool.size();
      d.emission_intensity = m.emission_intensity;
      pool.insert(pool.end(), m.albedo.pixel_data.begin(), m.albedo.pixel_data.end());
      mats.push_back(d);
    }
    for (const Light& l : s.lights) {
      shb_light d{};
      d.type = (int32_t)l.type;
      for (int i = 0; i < 3; ++i) {
        d.position[i] = l.position[i];
        d.direction[i] = l.direction[i];
        d.color[i] = l.color[i];
      }
      d.intensity = l.intensity;
      d.cos_inner_cone = l.cos_inner_cone;
      d.cos_outer_cone = l.cos_outer_cone;
      d.geometry_index = l.geometry_index;
      d.material_index = l.material_index;
      lights.push_back(d);
    }
    desc.num_vertices = (uint32_t)(pos.size() / 3);
    desc.positions = pos.data();
    desc.normals = any_n ? nrm.data() : nullptr;
    desc.texture_uvs = any_t ? tuv.data() : nullptr;
    desc.lightmap_uvs = any_l ? luv.data() : nullptr;
    desc.tangents = any_g ? tan.data() : nullptr;
    desc.num_indices = (uint32_t)idx.size();
    desc.indices = idx.data();
    desc.num_geometries = (uint32_t)geos.size();
    desc.geometries = geos.data();
    desc.num_materials = (uint32_t)mats.size();
    desc.materials = mats.data();
    desc.texel_pool_size = pool.size();
    desc.texel_pool = pool.data();
    desc.num_lights = (uint32_t)lights.size();
    desc.lights = lights.data();
  }
};

inline bool upload(shb_ctx* c, const Scene& scene) {
  if (!c) return false;
  FlatScene f(scene);
  int rc = shb_upload_scene(c, &f.desc);
  if (rc != SHB_OK) log_error("shb_upload_scene", c, rc);
  return rc == SHB_OK;
}

inline shb_raster_cfg raster_cfg(const RasterConfig& r) {
  return shb_raster_cfg{r.width, r.height, r.supersample_scale};
}

inline SurfacePoint from_abi(const shb_surface_point& p) {
  SurfacePoint s;
  for (int i = 0; i < 3; ++i) {
    s.position[i] = p.position[i];
    s.normal[i] = p.normal[i];
  }
  for (int i = 0; i < 4; ++i) s.tangent[i] = p.tangent[i];
  s.material_id = p.material_id;
  return s;
}

inline shb_surface_point to_abi(const SurfacePoint& s) {
  shb_surface_point p{};
  for (int i = 0; i < 3; ++i) {
    p.position[i] = s.position[i];
    p.normal[i] = s.normal[i];
  }
  for (int i = 0; i < 4; ++i) p.tangent[i] = s.tangent[i];
  p.material_id = s.material_id;
  return p;
}

inline void unpack_sh(const std::vector<float>& f, std::vector<SHCoeffs>& out) {
  out.resize(f.size() / 27);
  for (size_t t = 0; t < out.size(); ++t)
    for (int k = 0; k < 9; ++k)
      for (int c = 0; c < 3; ++c) out[t].coeffs[k][c] = f[t * 27 + k * 3 + c];
}

inline std::vector<float> pack_sh(const std::vector<SHCoeffs>& p) {
  std::vector<float> f(p.size() * 27);
  for (size_t t = 0; t < p.size(); ++t)
    for (int k = 0; k < 9; ++k)
      for (int c = 0; c < 3; ++c) f[t * 27 + k * 3 + c] = p[t].coeffs[k][c];
  return f;
}

}  // namespace detail

// Chooses the GPU of the per-thread context (before its first use).
inline void SetDefaultDevice(int device) { detail::default_device() = device; }

// ---------------------------------------------------------- rasterizer.h
// Replaces RasterizeScene (src/rasterizer.cpp:68-211).
inline std::vector<SurfacePoint> RasterizeScene(const Scene& scene, const RasterConfig& config) {
  shb_ctx* c = detail::thread_context().get();
  if (!detail::upload(c, scene)) return {};
  const size_t n = (size_t)config.width * config.supersample_scale * config.height *
                   config.supersample_scale;
  std::vector<shb_surface_point> pts(n);
  shb_raster_cfg rc = detail::raster_cfg(config);
  int rv = shb_rasterize(c, &rc, pts.data());
  if (rv != SHB_OK) {
    detail::log_error("shb_rasterize", c, rv);
    return {};
  }
  std::vector<SurfacePoint> out(n);
  for (size_t i = 0; i < n; ++i) out[i] = detail::from_abi(pts[i]);
  return out;
}

// Replaces CreateValidityMask (src/rasterizer.cpp:213-220); host-side, as
// in the reference (the device form runs inside BakeSHLightMapDilated).
inline std::vector<uint8_t> CreateValidityMask(const std::vector<SurfacePoint>& points) {
  std::vector<uint8_t> m(points.size());
  for (size_t i = 0; i < points.size(); ++i) m[i] = points[i].material_id >= 0 ? 1 : 0;
  return m;
}

// Replaces DownsampleValidityMask (src/rasterizer.cpp:222-241).
inline std::vector<uint8_t> DownsampleValidityMask(const std::vector<uint8_t>& high_res_mask,
                                                   int width, int height, int scale) {
  shb_ctx* c = detail::thread_context().get();
  if (!c || high_res_mask.size() != (size_t)width * scale * height * scale) {
    std::fprintf(stderr, "shb_baker: DownsampleValidityMask: bad size\n");
    return {};
  }
  std::vector<uint8_t> out((size_t)width * height);
  int rc = shb_downsample_mask(c, width, height, scale, high_res_mask.data(), out.data());
  if (rc != SHB_OK) {
    detail::log_error("shb_downsample_mask", c, rc);
    return {};
  }
  return out;
}

// ---------------------------------------------------------------- baker.h
// Replaces BakeSHLightMap (src/baker.cpp:186-304): the scene's BVH is built
// and uploaded inside the call, as the reference builds its Embree scene.
inline BakeResult BakeSHLightMap(const Scene& scene, const std::vector<SurfacePoint>& surface_points,
                                 const RasterConfig& raster_config, const BakeConfig& config) {
  BakeResult r;
  shb_ctx* c = detail::thread_context().get();
  const int w = raster_config.width * raster_config.supersample_scale;
  const int h = raster_config.height * raster_config.supersample_scale;
  if (surface_points.size() != (size_t)w * h) {
    std::fprintf(stderr, "shb_baker: BakeSHLightMap: %zu surface points for a %dx%d grid\n",
                 surface_points.size(), w, h);
    return r;
  }
  if (!detail::upload(c, scene)) return r;
  std::vector<shb_surface_point> pts(surface_points.size());
  for (size_t i = 0; i < pts.size(); ++i) pts[i] = detail::to_abi(surface_points[i]);
  shb_raster_cfg rc = detail::raster_cfg(raster_config);
  int rv = shb_set_surface(c, &rc, pts.data());
  if (rv != SHB_OK) {
    detail::log_error("shb_set_surface", c, rv);
    return r;
  }
  std::vector<float> sh((size_t)w * h * 27), vis((size_t)w * h);
  shb_bake_cfg bc{config.samples, config.bounces, config.num_light_samples, 0, config.seed};
  rv = shb_bake(c, &bc, sh.data(), vis.data(), nullptr);
  if (rv != SHB_OK) {
    detail::log_error("shb_bake", c, rv);
    return r;
  }
  r.sh_texture.width = w;
  r.sh_texture.height = h;
  detail::unpack_sh(sh, r.sh_texture.pixels);
  r.environment_visibility_texture.width = w;
  r.environment_visibility_texture.height = h;
  r.environment_visibility_texture.channels = 1;
  r.environment_visibility_texture.pixel_data = std::move(vis);
  return r;
}

// Replaces DownsampleSHTexture (src/baker.cpp:306-313).
inline SHTexture DownsampleSHTexture(const SHTexture& input, int scale) {
  SHTexture out;
  shb_ctx* c = detail::thread_context().get();
  if (!c || scale <= 0) return out;
  out.width = input.width / scale;
  out.height = input.height / scale;
  std::vector<float> in = detail::pack_sh(input.pixels), o((size_t)out.width * out.height * 27);
  int rc = shb_downsample(c, out.width, out.height, scale, 27, in.data(), o.data());
  if (rc != SHB_OK) {
    detail::log_error("shb_downsample", c, rc);
    return SHTexture{};
  }
  detail::unpack_sh(o, out.pixels);
  return out;
}

// Replaces DownsampleEnvironmentVisibilityTexture (src/baker.cpp:315-323).
inline Texture32F DownsampleEnvironmentVisibilityTexture(const Texture32F& input, int scale) {
  Texture32F out;
  shb_ctx* c = detail::thread_context().get();
  if (!c || scale <= 0) return out;
  out.width = input.width / scale;
  out.height = input.height / scale;
  out.channels = input.channels;
  out.pixel_data.resize((size_t)out.width * out.height);
  int rc = shb_downsample(c, (int)out.width, (int)out.height, scale, 1, input.pixel_data.data(),
                          out.pixel_data.data());
  if (rc != SHB_OK) {
    detail::log_error("shb_downsample", c, rc);
    return Texture32F{};
  }
  return out;
}

// ------------------------------------------------------------- dilation.h
// Replaces Dilate<T> (src/dilation.cpp:7-73) for T = SHCoeffs and float:
// pixels and mask are modified in place.
template <typename T>
void Dilate(int width, int height, std::vector<T>& pixels, std::vector<uint8_t>& mask,
            int passes = 1);

template <>
inline void Dilate<SHCoeffs>(int width, int height, std::vector<SHCoeffs>& pixels,
                             std::vector<uint8_t>& mask, int passes) {
  shb_ctx* c = detail::thread_context().get();
  std::vector<float> f = detail::pack_sh(pixels);
  int rc = shb_dilate(c, width, height, 27, f.data(), mask.data(), passes);
  if (rc != SHB_OK) {
    detail::log_error("shb_dilate", c, rc);
    return;
  }
  detail::unpack_sh(f, pixels);
}

template <>
inline void Dilate<float>(int width, int height, std::vector<float>& pixels,
                          std::vector<uint8_t>& mask, int passes) {
  shb_ctx* c = detail::thread_context().get();
  int rc = shb_dilate(c, width, height, 1, pixels.data(), mask.data(), passes);
  if (rc != SHB_OK) detail::log_error("shb_dilate", c, rc);
}

// --------------------------------------------------------------- saver.h
// Replaces SaveMode / SaveSHLightMap (src/saver.h:19-47, src/saver.cpp:
// 103-422): OpenEXR, HALF channels, ZIP. kCombined: one file of 28 channels
// (L0.R .. L22.B, EnvVisibility); kSplitChannels: nine files path_L0 ..
// path_L22 (RGB; L0 also carries visibility in A); kLuminancePacked: three
// RGBA files path_packed_0..2 (L0 RGB + visibility, then the Rec. 709 luma
// of coefficients 1-4 and 5-8). Host code: needs no GPU.
enum class SaveMode { kCombined, kSplitChannels, kLuminancePacked };

inline bool SaveSHLightMap(const SHTexture& sh_texture,
                           const Texture32F& environment_visibility_texture,
                           const std::filesystem::path& path,
                           SaveMode mode = SaveMode::kCombined) {
  if (sh_texture.pixels.empty() || sh_texture.width <= 0 || sh_texture.height <= 0 ||
      sh_texture.pixels.size() < (size_t)sh_texture.width * sh_texture.height) {
    std::fprintf(stderr, "shb_baker: invalid SHTexture dimensions or empty pixels\n");
    return false;
  }
  const size_t n = (size_t)sh_texture.width * sh_texture.height;
  std::vector<float> sh = detail::pack_sh(sh_texture.pixels);
  // missing visibility pixels are written as 0 (src/saver.cpp:166-172)
  std::vector<float> vis(n, 0.0f);
  const auto& v = environment_visibility_texture.pixel_data;
  std::copy(v.begin(), v.begin() + std::min(v.size(), n), vis.begin());
  const int32_t m = mode == SaveMode::kCombined ? SHB_SAVE_COMBINED
                    : mode == SaveMode::kSplitChannels ? SHB_SAVE_SPLIT
                                                       : SHB_SAVE_LUMINANCE_PACKED;
  char err[512] = {0};
  int rc = shb_save_lightmap(path.string().c_str(), sh_texture.width, sh_texture.height,
                             sh.data(), vis.data(), m, err, sizeof(err));
  if (rc != SHB_OK) {
    std::fprintf(stderr, "shb_baker: SaveSHLightMap failed: %s\n", err);
    return false;
  }
  return true;
}

// ------------------------------------------------------------- scene.h
// Replaces ProjectEnvironmentToSH (src/scene.cpp:152-158): a LatLong texture
// (channels >= 3) is projected on the GPU (per-thread context), the Preetham
// sky on the host. On failure logs and returns zero coefficients.
inline SHCoeffs ProjectEnvironmentToSH(const Environment& env) {
  SHCoeffs out;
  float f[27] = {0};
  int rc;
  if (env.type == Environment::Type::Texture) {
    shb_ctx* c = detail::thread_context().get();
    if (!c) return out;
    const Texture32F& t = env.texture;
    if ((size_t)t.width * t.height * t.channels > t.pixel_data.size()) {
      std::fprintf(stderr, "shb_baker: environment texture smaller than its size\n");
      return out;
    }
    rc = shb_env_sh_latlong(c, (int32_t)t.width, (int32_t)t.height, (int32_t)t.channels,
                            t.pixel_data.data(), f, nullptr);
    if (rc != SHB_OK) {
      detail::log_error("shb_env_sh_latlong", c, rc);
      return out;
    }
  } else {
    const float s[3] = {env.sun_direction.x(), env.sun_direction.y(), env.sun_direction.z()};
    rc = shb_env_sh_preetham(s, f);
    if (rc != SHB_OK) {
      std::fprintf(stderr, "shb_baker: shb_env_sh_preetham failed (%d)\n", rc);
      return out;
    }
  }
  for (int i = 0; i < 9; ++i)
    for (int k = 0; k < 3; ++k) out.coeffs[i][k] = f[3 * i + k];
  return out;
}

// ------------------------------------------------------------ loader.h
// Replaces LoadScene (src/loader.h, src/loader.cpp:113-844) over
// shb_load_gltf: the same Scene, with the environment's SH projected as the
// reference does at load (a skybox texture on the GPU of the per-thread
// context, a Preetham sky on the host). std::nullopt (and a log line) when
// the file cannot be loaded.
inline std::optional<Scene> LoadScene(const std::filesystem::path& gltf_file) {
  shb_gltf* g = nullptr;
  char err[1024] = {0};
  if (shb_load_gltf(gltf_file.string().c_str(), &g, err, sizeof(err)) != SHB_OK) {
    std::fprintf(stderr, "shb_baker: LoadScene(%s): %s\n", gltf_file.string().c_str(), err);
    return std::nullopt;
  }
  const shb_scene_desc& d = *shb_gltf_scene(g);
  Scene s;
  for (uint32_t gi = 0; gi < d.num_geometries; ++gi) {
    const shb_geometry& sg = d.geometries[gi];
    Geometry geo;
    for (uint32_t v = sg.vertex_offset; v < sg.vertex_offset + sg.vertex_count; ++v) {
      geo.vertices.emplace_back(d.positions[3 * v], d.positions[3 * v + 1], d.positions[3 * v + 2]);
      geo.normals.emplace_back(d.normals[3 * v], d.normals[3 * v + 1], d.normals[3 * v + 2]);
      geo.texture_uvs.emplace_back(d.texture_uvs[2 * v], d.texture_uvs[2 * v + 1]);
      if (sg.flags & SHB_GEOM_HAS_LIGHTMAP_UVS)
        geo.lightmap_uvs.emplace_back(d.lightmap_uvs[2 * v], d.lightmap_uvs[2 * v + 1]);
      geo.tangents.emplace_back(d.tangents[4 * v], d.tangents[4 * v + 1], d.tangents[4 * v + 2],
                                d.tangents[4 * v + 3]);
    }
    geo.indices.assign(d.indices + sg.index_offset, d.indices + sg.index_offset + sg.index_count);
    geo.material_id = sg.material_id;
    std::memcpy(geo.transform.m.data(), sg.transform, sizeof(sg.transform));
    s.geometries.push_back(std::move(geo));
  }
  for (uint32_t mi = 0; mi < d.num_materials; ++mi) {
    const shb_material& sm = d.materials[mi];
    Material m;
    const char* name = shb_gltf_material_name(g, mi);
    m.name = name ? name : "";
    m.albedo.width = sm.albedo_width;
    m.albedo.height = sm.albedo_height;
    m.albedo.channels = sm.albedo_channels;
    const size_t n = (size_t)sm.albedo_width * sm.albedo_height * sm.albedo_channels;
    m.albedo.pixel_data.assign(d.texel_pool + sm.albedo_offset, d.texel_pool + sm.albedo_offset + n);
    if (const char* ap = shb_gltf_albedo_path(g, mi)) m.albedo.file_path = std::filesystem::path(ap);
    m.emission_intensity = sm.emission_intensity;
    s.materials.push_back(std::move(m));
  }
  for (uint32_t li = 0; li < d.num_lights; ++li) {
    const shb_light& sl = d.lights[li];
    Light l;
    l.type = static_cast<Light::Type>(sl.type);
    l.position = Vector3f(sl.position[0], sl.position[1], sl.position[2]);
    l.direction = Vector3f(sl.direction[0], sl.direction[1], sl.direction[2]);
    l.color = Vector3f(sl.color[0], sl.color[1], sl.color[2]);
    l.intensity = sl.intensity;
    l.cos_inner_cone = sl.cos_inner_cone;
    l.cos_outer_cone = sl.cos_outer_cone;
    l.geometry_index = sl.geometry_index;
    l.material_index = sl.material_index;
    s.lights.push_back(l);
  }
  shb_environment_desc e{};
  shb_gltf_environment(g, &e);
  if (e.type != SHB_ENV_NONE) {
    Environment env;
    if (e.type == SHB_ENV_TEXTURE) {
      env.type = Environment::Type::Texture;
      env.texture.width = (uint32_t)e.width;
      env.texture.height = (uint32_t)e.height;
      env.texture.channels = 3;
      env.texture.pixel_data.assign(e.pixels, e.pixels + (size_t)e.width * e.height * 3);
      if (e.path) env.texture.file_path = std::filesystem::path(e.path);
    } else {
      env.type = Environment::Type::Preetham;
      env.sun_direction = Vector3f(e.sun_direction[0], e.sun_direction[1], e.sun_direction[2]);
      env.sun_intensity = e.sun_intensity;
    }
    env.sh_coeffs = ProjectEnvironmentToSH(env);
    s.environment = std::move(env);
  }
  shb_gltf_free(g);
  return s;
}

// ------------------------------------------------------------- atlas.h
namespace atlas_internal {
// Replaces atlas_internal::CalculateGeometryScales (src/atlas.h,
// src/atlas.cpp:22-95).
inline std::vector<float> CalculateGeometryScales(const std::vector<Geometry>& geometries,
                                                  const std::vector<Material>& materials,
                                                  float density_multiplier) {
  // only texture UVs and albedo sizes matter: a description of just those
  std::vector<float> pos, uv;
  std::vector<shb_geometry> geos;
  for (const Geometry& g : geometries) {
    shb_geometry d{};
    d.vertex_offset = (uint32_t)(uv.size() / 2);
    d.vertex_count = (uint32_t)g.texture_uvs.size();
    d.material_id = g.material_id;
    d.flags = SHB_GEOM_HAS_TEXTURE_UVS;
    for (const Vector2f& t : g.texture_uvs) {
      uv.push_back(t.x());
      uv.push_back(t.y());
      pos.insert(pos.end(), 3, 0.0f);
    }
    geos.push_back(d);
  }
  std::vector<shb_material> mats;
  for (const Material& m : materials) {
    shb_material d{};
    d.albedo_width = m.albedo.width;
    d.albedo_height = m.albedo.height;
    mats.push_back(d);
  }
  shb_scene_desc desc{};
  desc.num_vertices = (uint32_t)(pos.size() / 3);
  desc.positions = pos.data();
  desc.texture_uvs = uv.data();
  desc.num_geometries = (uint32_t)geos.size();
  desc.geometries = geos.data();
  desc.num_materials = (uint32_t)mats.size();
  desc.materials = mats.data();
  std::vector<float> out(geometries.size());
  if (shb_atlas_geometry_scales(&desc, density_multiplier, out.data()) != SHB_OK) {
    std::fprintf(stderr, "shb_baker: CalculateGeometryScales: bad materials or texture UVs\n");
    return {};
  }
  return out;
}
}  // namespace atlas_internal

struct AtlasResult {
  std::vector<Geometry> geometries;
  int width = 0;
  int height = 0;
};

// Replaces CreateAtlasGeometries (src/atlas.h, src/atlas.cpp:97-250): the
// geometries with split vertices and lightmap UVs in one
// target_resolution^2 atlas; std::nullopt when the charts cannot fit.
inline std::optional<AtlasResult> CreateAtlasGeometries(const Scene& scene,
                                                        unsigned target_resolution,
                                                        unsigned padding,
                                                        float density_multiplier = 1.0f) {
  if (scene.geometries.empty()) return std::nullopt;
  detail::FlatScene f(scene);
  shb_atlas* a = nullptr;
  char err[1024] = {0};
  if (shb_atlas_create(&f.desc, target_resolution, padding, density_multiplier, &a, err,
                       sizeof(err)) != SHB_OK) {
    std::fprintf(stderr, "shb_baker: CreateAtlasGeometries: %s\n", err);
    return std::nullopt;
  }
  const shb_scene_desc& d = *shb_atlas_scene(a);
  AtlasResult r;
  shb_atlas_info(a, &r.width, &r.height, nullptr, nullptr, nullptr, nullptr);
  for (uint32_t gi = 0; gi < d.num_geometries; ++gi) {
    const shb_geometry& sg = d.geometries[gi];
    Geometry g;
    for (uint32_t v = sg.vertex_offset; v < sg.vertex_offset + sg.vertex_count; ++v) {
      g.vertices.emplace_back(d.positions[3 * v], d.positions[3 * v + 1], d.positions[3 * v + 2]);
      g.normals.emplace_back(d.normals[3 * v], d.normals[3 * v + 1], d.normals[3 * v + 2]);
      g.texture_uvs.emplace_back(d.texture_uvs[2 * v], d.texture_uvs[2 * v + 1]);
      g.lightmap_uvs.emplace_back(d.lightmap_uvs[2 * v], d.lightmap_uvs[2 * v + 1]);
      g.tangents.emplace_back(d.tangents[4 * v], d.tangents[4 * v + 1], d.tangents[4 * v + 2],
                              d.tangents[4 * v + 3]);
    }
    g.indices.assign(d.indices + sg.index_offset, d.indices + sg.index_offset + sg.index_count);
    g.material_id = sg.material_id;
    std::memcpy(g.transform.m.data(), sg.transform, sizeof(sg.transform));
    r.geometries.push_back(std::move(g));
  }
  shb_atlas_free(a);
  return r;
}

// ------------------------------------------------------------ occlusion.h
// An uploaded scene with its BVH: the counterpart of the RTCScene that
// BuildBVH (src/scene.cpp:209-262) returns. Owns its own GPU context.
class BvhScene {
 public:
  explicit BvhScene(const Scene& scene, int device = 0) : ctx_(device) {
    ok_ = detail::upload(ctx_.get(), scene);
  }
  bool valid() const { return ok_; }
  shb_ctx* context() const { return ctx_.get(); }

 private:
  detail::Context ctx_;
  bool ok_ = false;
};

// Many rays at once (one kernel launch); the batch form of FindOcclusion.
inline std::vector<std::optional<Occlusion>> FindOcclusions(const BvhScene& scene,
                                                            const std::vector<Ray>& rays) {
  std::vector<std::optional<Occlusion>> out(rays.size());
  if (!scene.valid()) {
    std::fprintf(stderr, "shb_baker: invalid scene provided to FindOcclusion\n");
    return out;
  }
  std::vector<shb_ray> r(rays.size());
  for (size_t i = 0; i < rays.size(); ++i) {
    for (int k = 0; k < 3; ++k) {
      r[i].origin[k] = rays[i].origin[k];
      r[i].direction[k] = rays[i].direction[k];
    }
    r[i].tnear = rays[i].tnear;
    r[i].tfar = rays[i].tfar;
  }
  std::vector<shb_hit> h(rays.size());
  int rc = shb_trace_rays(scene.context(), r.data(), (uint32_t)r.size(), 0, h.data(), nullptr,
                          nullptr);
  if (rc != SHB_OK) {
    detail::log_error("shb_trace_rays", scene.context(), rc);
    return out;
  }
  for (size_t i = 0; i < rays.size(); ++i) {
    if (h[i].prim_id < 0) continue;
    Occlusion o;
    for (int k = 0; k < 3; ++k) {
      o.position[k] = h[i].position[k];
      o.normal[k] = h[i].normal[k];
    }
    o.uv[0] = h[i].uv[0];
    o.uv[1] = h[i].uv[1];
    o.material_id = h[i].material_id;
    out[i] = o;
  }
  return out;
}

// Replaces FindOcclusion (src/occlusion.cpp:9-107).
inline std::optional<Occlusion> FindOcclusion(const BvhScene& scene, const Ray& ray) {
  return FindOcclusions(scene, std::vector<Ray>{ray})[0];
}

}  // namespace shb_baker

#endif  // SHB_BAKER_HPP_
