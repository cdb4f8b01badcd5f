// The reference's own unit tests for the bake path (src/baker_test.cpp,
// src/occlusion_test.cpp, src/rasterizer_test.cpp, src/dilation_test.cpp,
// and the lightmap cases of src/saver_test.cpp),
// restated against the C++ mirror include/shb_baker.hpp: same scenes, same
// expectations, so a reference user sees the same contract hold on the GPU.
// GTest is not available here; the small harness below keeps the
// TEST / EXPECT_* spelling. Without a GPU only the HOST_TEST cases run and
// the exit code is 77 when they pass.
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <functional>
#include <string>
#include <vector>

#include "../../include/shb_baker.hpp"

namespace {
struct Case {
  const char* name;
  std::function<void()> fn;
  bool host_only;
};
std::vector<Case>& registry() {
  static std::vector<Case> r;
  return r;
}
struct Reg {
  Reg(const char* n, std::function<void()> f, bool host = false) {
    registry().push_back({n, std::move(f), host});
  }
};
int g_failures = 0;
bool g_case_failed = false;
void report(const char* file, int line, const std::string& what) {
  std::printf("  %s:%d: %s\n", file, line, what.c_str());
  g_case_failed = true;
}
}  // namespace

#define TEST(suite, name)                                                  \
  static void suite##_##name();                                            \
  static Reg reg_##suite##_##name(#suite "." #name, suite##_##name);       \
  static void suite##_##name()
#define HOST_TEST(suite, name)                                                \
  static void suite##_##name();                                               \
  static Reg reg_##suite##_##name(#suite "." #name, suite##_##name, true);    \
  static void suite##_##name()
#define CHECK_OP(a, op, b, fatal)                                                         \
  do {                                                                                    \
    auto va = (a);                                                                        \
    auto vb = (b);                                                                        \
    if (!(va op vb)) {                                                                    \
      report(__FILE__, __LINE__,                                                          \
             std::string(#a " " #op " " #b " failed: ") + std::to_string((double)va) +   \
                 " vs " + std::to_string((double)vb));                                    \
      if (fatal) return;                                                                  \
    }                                                                                     \
  } while (0)
#define EXPECT_EQ(a, b) CHECK_OP(a, ==, b, false)
#define ASSERT_EQ(a, b) CHECK_OP(a, ==, b, true)
#define EXPECT_GT(a, b) CHECK_OP(a, >, b, false)
#define EXPECT_GE(a, b) CHECK_OP(a, >=, b, false)
#define EXPECT_NEAR(a, b, tol) CHECK_OP(std::fabs((double)(a) - (double)(b)), <=, (double)(tol), false)
#define EXPECT_FLOAT_EQ(a, b) \
  CHECK_OP(std::fabs((double)(a) - (double)(b)), <=, 4.0 * 1.1920929e-7 * std::fabs((double)(b)), false)
#define EXPECT_TRUE(c) CHECK_OP((bool)(c), ==, true, false)
#define EXPECT_FALSE(c) CHECK_OP((bool)(c), ==, false, false)
#define ASSERT_TRUE(c) CHECK_OP((bool)(c), ==, true, true)

using namespace shb_baker;

static Geometry unit_quad(bool tangents) {
  Geometry quad;
  quad.material_id = 0;
  quad.vertices = {{-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0}};
  quad.normals = {{0, 0, 1}, {0, 0, 1}, {0, 0, 1}, {0, 0, 1}};
  quad.texture_uvs = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
  quad.lightmap_uvs = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
  if (tangents) quad.tangents = {{1, 0, 0, 1}, {1, 0, 0, 1}, {1, 0, 0, 1}, {1, 0, 0, 1}};
  quad.indices = {0, 1, 2, 0, 2, 3};
  return quad;
}

// baker_test.cpp:10-65
TEST(BakerTest, BakeSimpleQuad) {
  Scene scene;
  scene.geometries.push_back(unit_quad(false));
  Material mat;
  mat.name = "white";
  mat.albedo.pixel_data = {255, 255, 255, 255};
  mat.albedo.width = mat.albedo.height = 1;
  mat.albedo.channels = 4;
  scene.materials.push_back(mat);
  Light light;
  light.type = Light::Type::Directional;
  light.direction = Vector3f(0, 0, -1);
  scene.lights.push_back(light);

  std::vector<SurfacePoint> surface_points(16);
  RasterConfig raster_config;
  raster_config.width = raster_config.height = 4;
  surface_points[5].material_id = 0;
  surface_points[5].position = Vector3f(-0.25f, -0.25f, 0.0f);
  surface_points[5].normal = Vector3f(0.0f, 0.0f, 1.0f);
  surface_points[5].tangent = Vector4f(1.0f, 0.0f, 0.0f, 1.0f);
  BakeConfig config;
  config.samples = 64;
  config.bounces = 0;
  BakeResult output = BakeSHLightMap(scene, surface_points, raster_config, config);
  ASSERT_EQ(output.sh_texture.pixels.size(), (size_t)16);
  SHCoeffs result = output.sh_texture.pixels[5];
  EXPECT_GT(result.coeffs[0].x(), 0.1f);
  // closed form: every sample adds Y(0,0,1), 64 x (1/64)
  EXPECT_NEAR(result.coeffs[0].x(), 0.282095f, 1e-6f);
  EXPECT_NEAR(result.coeffs[2].y(), 0.488603f, 1e-6f);
  EXPECT_NEAR(result.coeffs[6].z(), 0.630784f, 1e-6f);
  EXPECT_EQ(output.environment_visibility_texture.pixel_data[5], 1.0f);
  EXPECT_EQ(output.sh_texture.pixels[0].coeffs[0].x(), 0.0f);
}

// baker_test.cpp:67-88
TEST(BakerTest, DownsampleSHTexture) {
  SHTexture input;
  input.width = input.height = 2;
  input.pixels.resize(4);
  for (auto& p : input.pixels)
    for (auto& c : p.coeffs) c = Vector3f(1.0f, 1.0f, 1.0f);
  SHTexture output = DownsampleSHTexture(input, 2);
  EXPECT_EQ(output.width, 1);
  EXPECT_EQ(output.height, 1);
  ASSERT_EQ(output.pixels.size(), (size_t)1);
  for (int c = 0; c < 9; ++c) EXPECT_FLOAT_EQ(output.pixels[0].coeffs[c].x(), 1.0f);
}

// occlusion_test.cpp:18-62
TEST(OcclusionTest, TriangleIntersection) {
  Scene scene;
  Geometry geo;
  geo.vertices = {{-1.0f, -1.0f, 0.0f}, {1.0f, -1.0f, 0.0f}, {0.0f, 1.0f, 0.0f}};
  geo.indices = {0, 1, 2};
  scene.geometries.push_back(geo);
  BvhScene bvh(scene);
  ASSERT_TRUE(bvh.valid());
  Ray ray_hit;
  ray_hit.origin = Vector3f(0.0f, 0.0f, 1.0f);
  ray_hit.direction = Vector3f(0.0f, 0.0f, -1.0f);
  auto hit = FindOcclusion(bvh, ray_hit);
  ASSERT_TRUE(hit.has_value());
  EXPECT_NEAR(hit->position.z(), 0.0f, 1e-4f);
  EXPECT_NEAR(hit->normal.z(), 1.0f, 1e-4f);
  Ray ray_miss;
  ray_miss.origin = Vector3f(2.0f, 0.0f, 1.0f);
  ray_miss.direction = Vector3f(0.0f, 0.0f, -1.0f);
  EXPECT_FALSE(FindOcclusion(bvh, ray_miss));
  Ray ray_away;
  ray_away.origin = Vector3f(0.0f, 0.0f, 1.0f);
  ray_away.direction = Vector3f(0.0f, 0.0f, 1.0f);
  EXPECT_FALSE(FindOcclusion(bvh, ray_away));
}

// occlusion_test.cpp:64-115
TEST(OcclusionTest, RayClipping) {
  Scene scene;
  Geometry geo;
  geo.vertices = {{-1.0f, -1.0f, 0.0f}, {1.0f, -1.0f, 0.0f}, {0.0f, 1.0f, 0.0f}};
  geo.indices = {0, 1, 2};
  scene.geometries.push_back(geo);
  BvhScene bvh(scene);
  ASSERT_TRUE(bvh.valid());
  Ray ray;
  ray.origin = Vector3f(0.0f, 0.0f, 1.0f);
  ray.direction = Vector3f(0.0f, 0.0f, -1.0f);
  ray.tfar = 0.5f;  // stops before the triangle at t = 1
  EXPECT_FALSE(FindOcclusion(bvh, ray));
  ray.tnear = 1.2f;  // starts behind it
  ray.tfar = std::numeric_limits<float>::infinity();
  EXPECT_FALSE(FindOcclusion(bvh, ray));
  ray.tnear = 0.1f;
  ray.tfar = 10.0f;
  auto hit = FindOcclusion(bvh, ray);
  ASSERT_TRUE(hit.has_value());
  EXPECT_NEAR(hit->position.z(), 0.0f, 1e-4f);
}

// occlusion_test.cpp:117-184
TEST(OcclusionTest, MultiGeometryMaterialTest) {
  Scene scene;
  Geometry geo_a, geo_b;
  geo_a.vertices = {{-3, -1, 0}, {-1, -1, 0}, {-3, 1, 0}, {-1, 1, 0}};
  geo_a.indices = {0, 1, 2, 1, 3, 2};
  geo_a.material_id = 10;
  geo_b.vertices = {{1, -1, 0}, {3, -1, 0}, {1, 1, 0}, {3, 1, 0}};
  geo_b.indices = {0, 1, 2, 1, 3, 2};
  geo_b.material_id = 20;
  scene.geometries.push_back(geo_a);
  scene.geometries.push_back(geo_b);
  BvhScene bvh(scene);
  ASSERT_TRUE(bvh.valid());
  Ray ray;
  ray.origin = Vector3f(2.8f, 0.8f, 5.0f);
  ray.direction = Vector3f(0.0f, 0.0f, -1.0f);
  auto hit = FindOcclusion(bvh, ray);
  ASSERT_TRUE(hit.has_value());
  EXPECT_EQ(hit->material_id, 20);
  EXPECT_NEAR(hit->position.x(), 2.8f, 1e-4f);
  EXPECT_NEAR(hit->position.y(), 0.8f, 1e-4f);
  EXPECT_NEAR(hit->position.z(), 0.0f, 1e-4f);
}

// rasterizer_test.cpp:7-43
TEST(RasterizerTest, RasterizeQuad) {
  Scene scene;
  scene.geometries.push_back(unit_quad(true));
  RasterConfig config;
  config.width = config.height = 4;
  std::vector<SurfacePoint> result = RasterizeScene(scene, config);
  ASSERT_EQ(result.size(), (size_t)16);
  for (int i = 0; i < 16; ++i) EXPECT_GE(result[i].material_id, 0);
  EXPECT_NEAR(result[5].position.x(), -0.25f, 0.001f);
  EXPECT_NEAR(result[5].position.y(), -0.25f, 0.001f);
  EXPECT_NEAR(result[5].normal.z(), 1.0f, 0.001f);
}

// rasterizer_test.cpp:45-72
TEST(RasterizerTest, RasterizeQuadSupersampled) {
  Scene scene;
  scene.geometries.push_back(unit_quad(true));
  RasterConfig config;
  config.width = config.height = 4;
  config.supersample_scale = 2;
  std::vector<SurfacePoint> result = RasterizeScene(scene, config);
  ASSERT_EQ(result.size(), (size_t)64);
  for (int i = 0; i < 64; ++i) EXPECT_GE(result[i].material_id, 0);
}

// rasterizer_test.cpp:74-87
TEST(RasterizerTest, ValidityMask) {
  std::vector<SurfacePoint> points(3);
  points[0].material_id = 0;
  points[1].material_id = -1;
  points[2].material_id = 0;
  std::vector<uint8_t> mask = CreateValidityMask(points);
  ASSERT_EQ(mask.size(), (size_t)3);
  EXPECT_EQ(mask[0], 1);
  EXPECT_EQ(mask[1], 0);
  EXPECT_EQ(mask[2], 1);
}

// rasterizer_test.cpp:89-126
TEST(RasterizerTest, DownsampleValidityMask) {
  std::vector<uint8_t> points(16, 0);
  points[2] = 1;
  points[8] = points[9] = points[12] = points[13] = 1;
  points[15] = 1;
  std::vector<uint8_t> mask = DownsampleValidityMask(points, 2, 2, 2);
  ASSERT_EQ(mask.size(), (size_t)4);
  EXPECT_EQ(mask[0], 0);
  EXPECT_EQ(mask[1], 1);
  EXPECT_EQ(mask[2], 1);
  EXPECT_EQ(mask[3], 1);
}

// dilation_test.cpp:11-33
TEST(DilationTest, SimpleFill) {
  int w = 3, h = 3;
  std::vector<SHCoeffs> pixels(w * h);
  std::vector<uint8_t> mask(w * h, 0);
  int center = 1 * w + 1;
  mask[center] = 1;
  pixels[center].coeffs[0] = Vector3f(1.0f, 0.0f, 0.0f);
  Dilate(w, h, pixels, mask, 1);
  EXPECT_EQ(mask[1], 1);
  EXPECT_NEAR(pixels[1].coeffs[0].x(), 1.0f, 0.001f);
  EXPECT_EQ(mask[0], 1);
  EXPECT_NEAR(pixels[0].coeffs[0].x(), 1.0f, 0.001f);
}

// dilation_test.cpp:35-58
TEST(DilationTest, AverageNeighbors) {
  int w = 3, h = 3;
  std::vector<SHCoeffs> pixels(w * h);
  std::vector<uint8_t> mask(w * h, 0);
  int left = 1 * w + 0, right = 1 * w + 2, center = 1 * w + 1;
  mask[left] = 1;
  pixels[left].coeffs[0] = Vector3f(1.0f, 0.0f, 0.0f);
  mask[right] = 1;
  pixels[right].coeffs[0] = Vector3f(0.0f, 0.0f, 1.0f);
  Dilate(w, h, pixels, mask, 1);
  EXPECT_EQ(mask[center], 1);
  EXPECT_NEAR(pixels[center].coeffs[0].x(), 0.5f, 0.001f);
  EXPECT_NEAR(pixels[center].coeffs[0].z(), 0.5f, 0.001f);
}

// the reference's glTF fixtures, copied under tests/golden/gltf
static std::filesystem::path gltf_dir() {
  char buf[4096] = {0};
  ssize_t n = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
  std::filesystem::path exe = n > 0 ? std::filesystem::path(std::string(buf, n)) : "";
  return exe.parent_path().parent_path() / "golden" / "gltf";
}

// atlas_test.cpp:9-65
HOST_TEST(AtlasTest, SimpleQuad) {
  Geometry input_geo;
  input_geo.vertices = {{-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0}};
  input_geo.normals = {{0, 0, 1}, {0, 0, 1}, {0, 0, 1}, {0, 0, 1}};
  input_geo.tangents = {{1, 0, 0, 1}, {1, 0, 0, 1}, {1, 0, 0, 1}, {1, 0, 0, 1}};
  input_geo.texture_uvs = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
  input_geo.indices = {0, 1, 2, 2, 1, 3};
  Scene scene;
  scene.geometries = {input_geo};
  Material mat;
  mat.albedo.width = 256;
  mat.albedo.height = 256;
  scene.materials.push_back(mat);
  scene.geometries[0].material_id = 0;
  std::optional<AtlasResult> atlas_result = CreateAtlasGeometries(scene, 256, 2);
  ASSERT_TRUE(atlas_result.has_value());
  EXPECT_GE(atlas_result->width, 0);
  EXPECT_GE(atlas_result->height, 0);
  ASSERT_EQ(atlas_result->geometries.size(), 1u);
  const auto& out_geo = atlas_result->geometries[0];
  ASSERT_EQ(out_geo.lightmap_uvs.size(), out_geo.vertices.size());
  for (const auto& uv : out_geo.lightmap_uvs) {
    EXPECT_GE(uv.x(), 0.0f);
    EXPECT_TRUE(uv.x() <= 1.0f);
    EXPECT_GE(uv.y(), 0.0f);
    EXPECT_TRUE(uv.y() <= 1.0f);
  }
  EXPECT_EQ(out_geo.indices.size(), input_geo.indices.size());
  EXPECT_GE(out_geo.vertices.size(), input_geo.vertices.size());
}

// atlas_test.cpp:67-166
HOST_TEST(AtlasTest, CalculateGeometryScales) {
  std::vector<Geometry> g(1);
  std::vector<Material> m(1);
  g[0].material_id = 0;
  g[0].texture_uvs = {{0, 0}, {1, 0}, {0, 1}};
  g[0].vertices = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  m[0].albedo.width = 100;
  m[0].albedo.height = 100;
  auto s = atlas_internal::CalculateGeometryScales(g, m, 1.0f);
  ASSERT_EQ(s.size(), 1u);
  EXPECT_NEAR(s[0], 100.0f, 1e-4f);  // Basic
  g[0].texture_uvs = {{0, 0}, {2, 0}, {0, 2}};
  m[0].albedo.width = m[0].albedo.height = 10;
  s = atlas_internal::CalculateGeometryScales(g, m, 1.0f);
  ASSERT_EQ(s.size(), 1u);
  EXPECT_NEAR(s[0], 20.0f, 1e-4f);  // Tiling
  std::vector<Geometry> g3(3);
  std::vector<Material> m3(3);
  for (int i = 0; i < 3; ++i) {
    g3[i].material_id = i;
    g3[i].texture_uvs = {{0, 0}, {1, 1}};
    g3[i].vertices = {{0, 0, 0}, {0, 0, 0}};
    m3[i].albedo.width = m3[i].albedo.height = i == 2 ? 1000 : 10;
  }
  s = atlas_internal::CalculateGeometryScales(g3, m3, 1.0f);
  ASSERT_EQ(s.size(), 3u);
  EXPECT_NEAR(s[0], 10.0f, 1e-4f);  // Clamping
  EXPECT_NEAR(s[1], 10.0f, 1e-4f);
  EXPECT_NEAR(s[2], 50.0f, 1e-4f);
  g[0].texture_uvs = {{0, 0}, {1, 1}};
  m[0].albedo.width = m[0].albedo.height = 100;
  s = atlas_internal::CalculateGeometryScales(g, m, 2.0f);
  ASSERT_EQ(s.size(), 1u);
  EXPECT_NEAR(s[0], 200.0f, 1e-4f);  // DensityMultiplier
}

// loader_test.cpp:9-50
HOST_TEST(LoaderTest, LoadCube) {
  auto input_path = gltf_dir() / "cube" / "Cube.gltf";
  ASSERT_TRUE(std::filesystem::exists(input_path));
  std::optional<Scene> scene = LoadScene(input_path);
  ASSERT_TRUE(scene.has_value());
  ASSERT_EQ(scene->geometries.size(), 1u);
  const auto& geo = scene->geometries[0];
  EXPECT_EQ(geo.indices.size(), 36u);
  EXPECT_GT(geo.vertices.size(), 0u);
  EXPECT_EQ(geo.vertices.size(), geo.normals.size());
  EXPECT_EQ(geo.vertices.size(), geo.texture_uvs.size());
  ASSERT_EQ(scene->materials.size(), 1u);
  EXPECT_TRUE(scene->materials[0].name == "Cube");
  EXPECT_GT(scene->materials[0].albedo.width, 0u);
  EXPECT_GT(scene->materials[0].albedo.height, 0u);
  ASSERT_TRUE(scene->materials[0].albedo.file_path.has_value());
  EXPECT_TRUE(scene->materials[0].albedo.file_path->is_absolute());
}

// loader_test.cpp:52-55
HOST_TEST(LoaderTest, MissingFile) {
  std::optional<Scene> scene = LoadScene("non_existent.gltf");
  EXPECT_FALSE(scene.has_value());
}

// loader_test.cpp:57-82
HOST_TEST(LoaderTest, LoadBoxFallbackColor) {
  auto input_path = gltf_dir() / "box" / "scene.gltf";
  ASSERT_TRUE(std::filesystem::exists(input_path));
  std::optional<Scene> scene = LoadScene(input_path);
  ASSERT_TRUE(scene.has_value());
  ASSERT_EQ(scene->materials.size(), 1u);
  const auto& mat = scene->materials[0];
  EXPECT_TRUE(mat.name == "Red");
  EXPECT_EQ(mat.albedo.width, 1u);
  EXPECT_EQ(mat.albedo.height, 1u);
  EXPECT_EQ(mat.albedo.channels, 4u);
  ASSERT_EQ(mat.albedo.pixel_data.size(), 4u);
  EXPECT_NEAR(mat.albedo.pixel_data[0], 231, 1);
  EXPECT_EQ(mat.albedo.pixel_data[1], 0);
  EXPECT_EQ(mat.albedo.pixel_data[2], 0);
  EXPECT_EQ(mat.albedo.pixel_data[3], 255);
  // the box's sun gives a Preetham environment (loader.cpp:750-789)
  ASSERT_TRUE(scene->environment.has_value());
  EXPECT_TRUE(scene->environment->type == Environment::Type::Preetham);
  EXPECT_GT(scene->environment->sh_coeffs.coeffs[0].z(), 0.0f);
}

// scene_test.cpp:96-126
TEST(SceneTest, ProjectEnvironmentUniformWhite) {
  Texture32F tex;
  tex.width = 64;
  tex.height = 32;
  tex.channels = 3;
  tex.pixel_data.resize(tex.width * tex.height * 3, 1.0f);
  Environment env;
  env.type = Environment::Type::Texture;
  env.texture = tex;
  SHCoeffs coeffs = ProjectEnvironmentToSH(env);
  float expected_L0 = 0.282095f * 4.0f * M_PI;
  EXPECT_NEAR(coeffs.coeffs[0].x(), expected_L0, 0.1f);
  EXPECT_NEAR(coeffs.coeffs[0].y(), expected_L0, 0.1f);
  EXPECT_NEAR(coeffs.coeffs[0].z(), expected_L0, 0.1f);
  for (int i = 1; i < 9; ++i) EXPECT_NEAR(coeffs.coeffs[i].x(), 0.0f, 0.1f);
}

// Preetham sky for a zenith sun: brighter above, symmetric about +Y
HOST_TEST(SceneTest, ProjectEnvironmentPreethamZenith) {
  Environment env;
  env.type = Environment::Type::Preetham;
  env.sun_direction = Vector3f(0, 1, 0);
  SHCoeffs c = ProjectEnvironmentToSH(env);
  EXPECT_GT(c.coeffs[0].z(), 0.0f);
  EXPECT_GT(c.coeffs[1].z(), 0.0f);
  EXPECT_NEAR(c.coeffs[3].z(), 0.0f, 1e-4f);
  EXPECT_NEAR(c.coeffs[0].x() / c.coeffs[0].z(), 0.2f / 0.9f, 1e-5f);
}

// saver_test.cpp: IsEXR checks the 4-byte magic; also the version field
static bool is_exr(const std::filesystem::path& p) {
  std::ifstream f(p, std::ios::binary);
  unsigned char b[8] = {0};
  f.read(reinterpret_cast<char*>(b), 8);
  return f.gcount() == 8 && b[0] == 0x76 && b[1] == 0x2f && b[2] == 0x31 && b[3] == 0x01 &&
         b[4] == 2;
}

static std::filesystem::path scratch_dir() {
  auto d = std::filesystem::temp_directory_path() /
           ("shb_saver_test_" + std::to_string((long)getpid()));
  std::filesystem::create_directories(d);
  return d;
}

static void fill_env(Texture32F& env, float v) {
  env.width = env.height = 16;
  env.pixel_data.assign(16 * 16, v);
}

// saver_test.cpp:15-46
HOST_TEST(SaverTest, SaveCombinedImage) {
  SHTexture tex;
  tex.width = tex.height = 16;
  tex.pixels.resize(16 * 16);
  Texture32F env_tex;
  fill_env(env_tex, 1.0f);
  for (auto& sh : tex.pixels) sh.coeffs[0] = Vector3f(1.0f, 0.5f, 0.25f);
  auto path = scratch_dir() / "test_output.exr";
  EXPECT_TRUE(SaveSHLightMap(tex, env_tex, path, SaveMode::kCombined));
  EXPECT_TRUE(std::filesystem::exists(path));
  EXPECT_TRUE(is_exr(path));
  std::filesystem::remove_all(path.parent_path());
}

// saver_test.cpp:48-103
HOST_TEST(SaverTest, SaveSplitChannels) {
  SHTexture tex;
  tex.width = tex.height = 16;
  tex.pixels.resize(16 * 16);
  Texture32F env_tex;
  fill_env(env_tex, 0.5f);
  for (auto& sh : tex.pixels) {
    sh.coeffs[0] = Vector3f(1.0f, 0.0f, 0.0f);
    sh.coeffs[1] = Vector3f(0.0f, 1.0f, 0.0f);
  }
  auto dir = scratch_dir();
  EXPECT_TRUE(SaveSHLightMap(tex, env_tex, dir / "test_split.exr", SaveMode::kSplitChannels));
  const char* suffixes[] = {"L0", "L1m1", "L10", "L11", "L2m2", "L2m1", "L20", "L21", "L22"};
  for (const char* s : suffixes) {
    auto f = dir / (std::string("test_split_") + s + ".exr");
    EXPECT_TRUE(std::filesystem::exists(f));
    EXPECT_TRUE(is_exr(f));
  }
  EXPECT_FALSE(std::filesystem::exists(dir / "test_split_EnvVisibility.exr"));
  std::filesystem::remove_all(dir);
}

// saver_test.cpp:217-259
HOST_TEST(SaverTest, SavePackedLuminance) {
  SHTexture tex;
  tex.width = tex.height = 16;
  tex.pixels.resize(16 * 16);
  Texture32F env_tex;
  fill_env(env_tex, 0.2f);
  for (auto& sh : tex.pixels) {
    sh.coeffs[0] = Vector3f(1.0f, 0.5f, 0.25f);
    sh.coeffs[1] = Vector3f(0.5f, 0.5f, 0.5f);
  }
  auto dir = scratch_dir();
  EXPECT_TRUE(SaveSHLightMap(tex, env_tex, dir / "test_packed.exr", SaveMode::kLuminancePacked));
  for (int i = 0; i < 3; ++i) {
    auto f = dir / ("test_packed_packed_" + std::to_string(i) + ".exr");
    EXPECT_TRUE(std::filesystem::exists(f));
    EXPECT_TRUE(is_exr(f));
  }
  std::filesystem::remove_all(dir);
}

// empty texture: logged error, false (saver.cpp:405-412)
HOST_TEST(SaverTest, RejectsEmptyTexture) {
  SHTexture tex;
  Texture32F env_tex;
  EXPECT_FALSE(SaveSHLightMap(tex, env_tex, scratch_dir() / "x.exr"));
  std::filesystem::remove_all(scratch_dir());
}

int main() {
  shb_ctx* probe = nullptr;
  const bool gpu = shb_create(&probe, 0) == SHB_OK;
  if (gpu) shb_destroy(probe);
  else std::printf("no GPU (%s): host-only cases\n", shb_version());
  int passed = 0;
  for (const Case& c : registry()) {
    if (!gpu && !c.host_only) continue;
    g_case_failed = false;
    std::printf("[ RUN  ] %s\n", c.name);
    std::fflush(stdout);
    c.fn();
    std::printf("[ %s ] %s\n", g_case_failed ? "FAIL" : " OK ", c.name);
    std::fflush(stdout);
    if (g_case_failed) ++g_failures; else ++passed;
  }
  std::printf("%d passed, %d failed\n", passed, g_failures);
  if (g_failures) return 1;
  return gpu ? 0 : 77;
}
