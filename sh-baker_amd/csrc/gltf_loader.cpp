// glTF 2.0 loader (host), replacing LoadScene (src/loader.cpp:113-844): a
// .gltf (JSON + external or data-URI buffers) or .glb file becomes the flat
// scene description the bake uploads (shb_scene_desc), plus the scene's
// environment.
//
// Same rules as the reference: the default scene's node tree is walked with
// affine transforms (matrix, or T * R * S); every mesh primitive needs a
// material, POSITION, NORMAL and TEXCOORD_0 (else the whole load fails);
// TEXCOORD_1 becomes the lightmap UVs; triangles with a repeated index are
// dropped; missing tangents are generated (from UV gradients where the
// reference uses MikkTSpace, else an arbitrary frame orthogonal to the
// normal); a primitive whose material has KHR_materials_emissive_strength >
// 0 becomes one area light; KHR_lights_punctual lights take position and
// -Z direction from their node; albedo is the decoded base-colour texture
// (RGBA8) or a 1x1 texture of the base-colour factor in sRGB; the
// environment is extras.skybox (Radiance HDR, or an LDR image / 255) or a
// Preetham sky towards the brightest directional light.
//
// Differences: only base-colour images are decoded (the bake uses no other
// texture), so an unsupported normal / metallic-roughness image does not
// fail the load; normalized integer texture coordinates are converted (the
// reference reads every attribute as float); the rotation that orients a
// light is the polar factor of the node's linear part (Eigen's rotation()
// for every transform with a positive determinant).
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <new>

#include "../../include/shb.h"
#include "image_decode.h"
#include "json_min.h"

using shb::json::Value;

namespace {

struct M34 {  // row-major 3x4 affine, float like Eigen::Affine3f
  float m[3][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}};
  // linear = A.linear * B.linear; translation = A.linear * B.translation + A.translation
  M34 operator*(const M34& b) const {
    M34 r;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 4; ++j) {
        float s = m[i][0] * b.m[0][j];
        s += m[i][1] * b.m[1][j];
        s += m[i][2] * b.m[2][j];
        if (j == 3) s += m[i][3];
        r.m[i][j] = s;
      }
    return r;
  }
};

}  // namespace

struct shb_gltf {
  std::vector<float> pos, nrm, tuv, luv, tan;
  std::vector<uint32_t> idx;
  std::vector<shb_geometry> geos;
  std::vector<shb_material> mats;
  std::vector<std::string> mat_names, albedo_paths;
  std::vector<uint8_t> texels;
  std::vector<shb_light> lights;
  shb_scene_desc desc{};
  int env_type = SHB_ENV_NONE;
  float sun[3] = {0, 1, 0};
  float sun_intensity = 1.0f;
  int env_w = 0, env_h = 0;
  std::vector<float> env_rgb;
  std::string env_path;
};

namespace {

struct Ctx {
  const Value* root = nullptr;
  std::string dir;  // directory of the .gltf, with trailing '/'
  std::vector<std::vector<uint8_t>> buffers;
  std::vector<uint8_t> glb_bin;
  bool has_glb_bin = false;
  shb_gltf* out = nullptr;
  std::string err;
};

bool read_file(const std::string& path, std::vector<uint8_t>* data) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  f.seekg(0, std::ios::end);
  const std::streamoff n = f.tellg();
  if (n < 0) return false;
  f.seekg(0);
  data->resize((size_t)n);
  if (n) f.read(reinterpret_cast<char*>(data->data()), n);
  return (bool)f;
}

bool is_absolute(const std::string& p) { return !p.empty() && p[0] == '/'; }

std::string absolute(const std::string& p) {
  if (is_absolute(p)) return p;
  char buf[4096];
  if (getcwd(buf, sizeof(buf))) return std::string(buf) + "/" + p;
  return p;
}

// percent-decoding of relative URIs (glTF allows escaped characters)
std::string uri_decode(const std::string& s) {
  std::string o;
  for (size_t i = 0; i < s.size(); ++i) {
    if (s[i] == '%' && i + 2 < s.size() && std::isxdigit((unsigned char)s[i + 1]) &&
        std::isxdigit((unsigned char)s[i + 2])) {
      o.push_back((char)std::stoi(s.substr(i + 1, 2), nullptr, 16));
      i += 2;
    } else {
      o.push_back(s[i]);
    }
  }
  return o;
}

bool base64(const std::string& s, std::vector<uint8_t>* o) {
  auto val = [](char c) -> int {
    if (c >= 'A' && c <= 'Z') return c - 'A';
    if (c >= 'a' && c <= 'z') return c - 'a' + 26;
    if (c >= '0' && c <= '9') return c - '0' + 52;
    if (c == '+' || c == '-') return 62;
    if (c == '/' || c == '_') return 63;
    return -1;
  };
  uint32_t acc = 0;
  int bits = 0;
  for (char c : s) {
    if (c == '=' || c == '\n' || c == '\r' || c == ' ') continue;
    const int v = val(c);
    if (v < 0) return false;
    acc = (acc << 6) | (uint32_t)v;
    bits += 6;
    if (bits >= 8) {
      bits -= 8;
      o->push_back((uint8_t)(acc >> bits));
    }
  }
  return true;
}

// bytes of a URI: data:...;base64,... or a file relative to the .gltf
bool load_uri(Ctx& c, const std::string& uri, std::vector<uint8_t>* data, std::string* path) {
  if (uri.rfind("data:", 0) == 0) {
    const size_t comma = uri.find(',');
    if (comma == std::string::npos || uri.substr(0, comma).find(";base64") == std::string::npos) {
      c.err = "unsupported data URI";
      return false;
    }
    if (!base64(uri.substr(comma + 1), data)) {
      c.err = "bad base64 data URI";
      return false;
    }
    if (path) path->clear();
    return true;
  }
  const std::string dec = uri_decode(uri);
  const std::string p = is_absolute(dec) ? dec : c.dir + dec;
  if (!read_file(p, data)) {
    c.err = "cannot read " + p;
    return false;
  }
  if (path) *path = absolute(p);
  return true;
}

const Value* arr_at(const Value* root, const char* key, int i) {
  const Value* a = root->get(key);
  if (!a || !a->is_array() || i < 0 || (size_t)i >= a->size()) return nullptr;
  return &(*a)[i];
}

bool load_buffers(Ctx& c) {
  const Value* bufs = c.root->get("buffers");
  if (!bufs) return true;
  for (size_t i = 0; i < bufs->size(); ++i) {
    const Value& b = (*bufs)[i];
    std::vector<uint8_t> data;
    const Value* uri = b.get("uri");
    if (uri && uri->is_string()) {
      if (!load_uri(c, uri->str, &data, nullptr)) return false;
    } else if (i == 0 && c.has_glb_bin) {
      data = c.glb_bin;
    } else {
      c.err = "buffer " + std::to_string(i) + " has no data";
      return false;
    }
    const double len = shb::json::num(b, "byteLength", (double)data.size());
    if (len > (double)data.size()) {
      c.err = "buffer " + std::to_string(i) + " shorter than its byteLength";
      return false;
    }
    c.buffers.push_back(std::move(data));
  }
  return true;
}

int comp_size(int ct) {
  switch (ct) {
    case 5120: case 5121: return 1;
    case 5122: case 5123: return 2;
    case 5125: case 5126: return 4;
    default: return 0;
  }
}

int type_count(const std::string& t) {
  if (t == "SCALAR") return 1;
  if (t == "VEC2") return 2;
  if (t == "VEC3") return 3;
  if (t == "VEC4") return 4;
  if (t == "MAT4") return 16;
  return 0;
}

// Sizes and offsets of the file are JSON numbers (doubles). They are taken
// only when finite, integral, non-negative and at most kMaxFileSize, so the
// byte arithmetic below (sums of a few such values, products of a count with
// a stride of at most kMaxStride) stays far from 64-bit wrap-around.
constexpr double kMaxFileSize = 1099511627776.0;  // 2^40 bytes
constexpr size_t kMaxStride = 1u << 16;
constexpr size_t kMaxZeroElements = (size_t)1 << 27;  // accessors without a bufferView

bool size_field(Ctx& c, const Value& v, const char* key, size_t dflt, size_t* out,
                const std::string& what) {
  const Value* m = v.get(key);
  if (!m) {
    *out = dflt;
    return true;
  }
  const double x = m->is_number() ? m->num : -1.0;
  if (!(x >= 0.0) || !(x <= kMaxFileSize) || x != std::floor(x)) {
    c.err = what + " has an invalid " + key;
    return false;
  }
  *out = (size_t)x;
  return true;
}

// [offset, offset + length) of a bufferView, checked against its buffer.
bool view_range(Ctx& c, const Value& bv, size_t buf_size, size_t* off, size_t* len,
                const std::string& what) {
  if (!size_field(c, bv, "byteOffset", 0, off, what) ||
      !size_field(c, bv, "byteLength", 0, len, what))
    return false;
  if (*off > buf_size || *len > buf_size - *off) {
    c.err = what + " lies beyond its buffer";
    return false;
  }
  return true;
}

// accessor elements as doubles (ncomp per element); normalized integers are
// mapped to [0, 1] / [-1, 1] as the glTF specification defines
bool read_accessor(Ctx& c, int ai, int want_comps, std::vector<double>* out, size_t* count,
                   bool integer_ok = false) {
  const Value* a = arr_at(c.root, "accessors", ai);
  if (!a) {
    c.err = "accessor " + std::to_string(ai) + " out of range";
    return false;
  }
  if (a->has("sparse")) {
    c.err = "sparse accessors are not supported";
    return false;
  }
  const std::string what = "accessor " + std::to_string(ai);
  const int ct = shb::json::integer(*a, "componentType", 0);
  const int nc = type_count(shb::json::string(*a, "type"));
  const int cs = comp_size(ct);
  size_t n = 0;
  if (!size_field(c, *a, "count", 0, &n, what)) return false;
  if (!cs || nc < want_comps || (want_comps && nc != want_comps)) {
    c.err = "accessor " + std::to_string(ai) + " has an unexpected type";
    return false;
  }
  const bool normalized = a->get("normalized") && a->get("normalized")->b;
  if (ct != 5126 && !integer_ok && !normalized) {
    c.err = "accessor " + std::to_string(ai) + " is not FLOAT";
    return false;
  }
  const int bvi = shb::json::integer(*a, "bufferView", -1);
  if (bvi < 0) {  // all zeros
    if (n > kMaxZeroElements) {
      c.err = what + " has too many elements for an accessor without a bufferView";
      return false;
    }
    out->assign(n * nc, 0.0);
    *count = n;
    return true;
  }
  const Value* bv = arr_at(c.root, "bufferViews", bvi);
  if (!bv) {
    c.err = "bufferView out of range";
    return false;
  }
  const int bi = shb::json::integer(*bv, "buffer", -1);
  if (bi < 0 || (size_t)bi >= c.buffers.size()) {
    c.err = "buffer out of range";
    return false;
  }
  const std::vector<uint8_t>& buf = c.buffers[bi];
  const size_t elem = (size_t)cs * nc;
  size_t view_off = 0, view_len = 0, acc_off = 0, stride = 0;
  if (!view_range(c, *bv, buf.size(), &view_off, &view_len, "bufferView " + std::to_string(bvi)) ||
      !size_field(c, *a, "byteOffset", 0, &acc_off, what) ||
      !size_field(c, *bv, "byteStride", 0, &stride, "bufferView " + std::to_string(bvi)))
    return false;
  if (stride == 0) stride = elem;
  if (stride < elem || stride > kMaxStride) {
    c.err = "bufferView " + std::to_string(bvi) + " has a byteStride below its element size "
            "or above 65536";
    return false;
  }
  // elements [0, n) at acc_off + i * stride, each elem bytes, inside the view
  if (acc_off > view_len ||
      (n && (elem > view_len - acc_off || n - 1 > (view_len - acc_off - elem) / stride))) {
    c.err = what + " reads past its buffer view";
    return false;
  }
  const size_t base = view_off + acc_off;
  out->assign(n * nc, 0.0);
  *count = n;
  for (size_t i = 0; i < n; ++i) {
    const uint8_t* e = &buf[base + i * stride];
    for (int k = 0; k < nc; ++k) {
      const uint8_t* q = e + (size_t)k * cs;
      double v = 0;
      switch (ct) {
        case 5126: { float f; std::memcpy(&f, q, 4); v = f; break; }
        case 5121: v = q[0]; if (normalized) v /= 255.0; break;
        case 5120: v = (int8_t)q[0]; if (normalized) v = std::max(v / 127.0, -1.0); break;
        case 5123: { uint16_t u; std::memcpy(&u, q, 2); v = u; if (normalized) v /= 65535.0; break; }
        case 5122: { int16_t s; std::memcpy(&s, q, 2); v = s; if (normalized) v = std::max(v / 32767.0, -1.0); break; }
        case 5125: { uint32_t u; std::memcpy(&u, q, 4); v = u; break; }
      }
      (*out)[i * nc + k] = v;
    }
  }
  return true;
}

// Eigen's quaternion -> rotation matrix (x, y, z, w as stored by glTF),
// in float as Eigen::Quaternionf
void quat_matrix(float x, float y, float z, float w, float r[3][3]) {
  const float tx = 2 * x, ty = 2 * y, tz = 2 * z;
  const float twx = tx * w, twy = ty * w, twz = tz * w;
  const float txx = tx * x, txy = ty * x, txz = tz * x;
  const float tyy = ty * y, tyz = tz * y, tzz = tz * z;
  r[0][0] = 1 - (tyy + tzz); r[0][1] = txy - twz;       r[0][2] = txz + twy;
  r[1][0] = txy + twz;       r[1][1] = 1 - (txx + tzz); r[1][2] = tyz - twx;
  r[2][0] = txz - twy;       r[2][1] = tyz + twx;       r[2][2] = 1 - (txx + tyy);
}

// src/loader.cpp:113-142 NodeToTransform; arithmetic in float as Eigen::Affine3f
M34 node_transform(const Value& node) {
  M34 t;
  const Value* mat = node.get("matrix");
  if (mat && mat->is_array() && mat->size() == 16) {
    for (int col = 0; col < 4; ++col)
      for (int row = 0; row < 3; ++row) t.m[row][col] = (float)(*mat)[col * 4 + row].number_or(0);
    return t;
  }
  const Value* tr = node.get("translation");
  if (tr && tr->is_array() && tr->size() == 3)
    for (int i = 0; i < 3; ++i) t.m[i][3] = (float)(*tr)[i].number_or(0);
  float r[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  const Value* ro = node.get("rotation");
  if (ro && ro->is_array() && ro->size() == 4)
    quat_matrix((float)(*ro)[0].number_or(0), (float)(*ro)[1].number_or(0),
                (float)(*ro)[2].number_or(0), (float)(*ro)[3].number_or(1), r);
  float s[3] = {1, 1, 1};
  const Value* sc = node.get("scale");
  if (sc && sc->is_array() && sc->size() == 3)
    for (int i = 0; i < 3; ++i) s[i] = (float)(*sc)[i].number_or(1);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t.m[i][j] = r[i][j] * s[j];
  return t;
}

// orthogonal polar factor of the linear part (Newton: R <- (R + R^-T) / 2)
void polar_rotation(const M34& t, double R[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i][j] = t.m[i][j];
  for (int it = 0; it < 60; ++it) {
    const double det = R[0][0] * (R[1][1] * R[2][2] - R[1][2] * R[2][1]) -
                       R[0][1] * (R[1][0] * R[2][2] - R[1][2] * R[2][0]) +
                       R[0][2] * (R[1][0] * R[2][1] - R[1][1] * R[2][0]);
    if (!(std::fabs(det) > 1e-30)) return;
    double cof[3][3];  // R^-T = cofactor / det
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        cof[i][j] = (R[i1][j1] * R[i2][j2] - R[i1][j2] * R[i2][j1]) / det;
      }
    double diff = 0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const double v = 0.5 * (R[i][j] + cof[i][j]);
        diff = std::max(diff, std::fabs(v - R[i][j]));
        R[i][j] = v;
      }
    if (diff < 1e-15) break;
  }
}

uint8_t linear_to_srgb(float x) {  // src/colorspace.h:17-22
  if (x <= 0.0031308f) return (uint8_t)std::rint(x * 12.92f * 255.0f);
  return (uint8_t)std::rint((1.055f * std::pow(x, 1.0f / 2.4f) - 0.055f) * 255.0f);
}

bool load_image(Ctx& c, int img_index, shb::Image8* img, std::string* path) {
  const Value* im = arr_at(c.root, "images", img_index);
  if (!im) {
    c.err = "image out of range";
    return false;
  }
  std::vector<uint8_t> bytes;
  const Value* uri = im->get("uri");
  if (uri && uri->is_string()) {
    if (!load_uri(c, uri->str, &bytes, path)) return false;
  } else {
    const int bvi = shb::json::integer(*im, "bufferView", -1);
    const Value* bv = arr_at(c.root, "bufferViews", bvi);
    if (!bv) {
      c.err = "image without uri or bufferView";
      return false;
    }
    const int bi = shb::json::integer(*bv, "buffer", -1);
    size_t off = 0, len = 0;
    if (bi < 0 || (size_t)bi >= c.buffers.size()) {
      c.err = "image bufferView out of range";
      return false;
    }
    if (!view_range(c, *bv, c.buffers[bi].size(), &off, &len, "image bufferView")) return false;
    bytes.assign(c.buffers[bi].begin() + off, c.buffers[bi].begin() + off + len);
    path->clear();
  }
  std::string e;
  if (!shb::decode_image8(bytes.data(), bytes.size(), img, &e)) {
    c.err = "image " + std::to_string(img_index) + ": " + e;
    return false;
  }
  return true;
}

// src/loader.cpp:440-543 ProcessMaterials (albedo and emission)
bool process_materials(Ctx& c) {
  shb_gltf& o = *c.out;
  const Value* mats = c.root->get("materials");
  auto add = [&](const std::string& name, int w, int h, const uint8_t* px, float emission,
                 const std::string& path) {
    shb_material m{};
    m.albedo_width = (uint32_t)w;
    m.albedo_height = (uint32_t)h;
    m.albedo_channels = 4;
    m.albedo_offset = o.texels.size();
    m.emission_intensity = emission;
    o.texels.insert(o.texels.end(), px, px + (size_t)w * h * 4);
    o.mats.push_back(m);
    o.mat_names.push_back(name);
    o.albedo_paths.push_back(path);
  };
  if (!mats || !mats->is_array() || mats->size() == 0) {
    const uint8_t white[4] = {255, 255, 255, 255};
    add("default", 1, 1, white, 0.0f, "");
    return true;
  }
  for (size_t i = 0; i < mats->size(); ++i) {
    const Value& gm = (*mats)[i];
    float emission = 0.0f;
    const Value* ext = gm.get("extensions");
    const Value* es = ext ? ext->get("KHR_materials_emissive_strength") : nullptr;
    if (es && es->get("emissiveStrength")) emission = (float)es->get("emissiveStrength")->number_or(0);
    const Value* pbr = gm.get("pbrMetallicRoughness");
    const Value* bct = pbr ? pbr->get("baseColorTexture") : nullptr;
    int tex = bct ? shb::json::integer(*bct, "index", -1) : -1;
    const Value* t = arr_at(c.root, "textures", tex);
    const int src = t ? shb::json::integer(*t, "source", -1) : -1;
    if (t && src >= 0 && arr_at(c.root, "images", src)) {
      shb::Image8 img;
      std::string path;
      if (!load_image(c, src, &img, &path)) return false;
      add(shb::json::string(gm, "name"), img.width, img.height, img.rgba.data(), emission, path);
      continue;
    }
    // 1x1 texture from baseColorFactor (linear -> sRGB); default factor 1
    double f[4] = {1, 1, 1, 1};
    const Value* bcf = pbr ? pbr->get("baseColorFactor") : nullptr;
    if (bcf && bcf->is_array() && bcf->size() == 4)
      for (int k = 0; k < 4; ++k) f[k] = (*bcf)[k].number_or(1);
    const uint8_t px[4] = {linear_to_srgb((float)f[0]), linear_to_srgb((float)f[1]),
                           linear_to_srgb((float)f[2]),
                           (uint8_t)std::rint(std::min(1.0, std::max(0.0, f[3])) * 255.0)};
    add(shb::json::string(gm, "name"), 1, 1, px, emission, "");
  }
  return true;
}

// UV-gradient tangents (accumulated per vertex, Gram-Schmidt against the
// normal, handedness from the bitangent); a vertex without a usable UV
// gradient gets the reference's geometric fallback frame
void generate_tangents(const std::vector<float>& P, const std::vector<float>& N,
                       const std::vector<float>& UV, const std::vector<uint32_t>& I, size_t nv,
                       std::vector<float>* T) {
  std::vector<double> t(nv * 3, 0.0), b(nv * 3, 0.0);
  for (size_t f = 0; f + 2 < I.size(); f += 3) {
    const uint32_t i0 = I[f], i1 = I[f + 1], i2 = I[f + 2];
    if (i0 >= nv || i1 >= nv || i2 >= nv) continue;
    double e1[3], e2[3];
    for (int k = 0; k < 3; ++k) {
      e1[k] = (double)P[3 * i1 + k] - P[3 * i0 + k];
      e2[k] = (double)P[3 * i2 + k] - P[3 * i0 + k];
    }
    const double du1 = (double)UV[2 * i1] - UV[2 * i0], dv1 = (double)UV[2 * i1 + 1] - UV[2 * i0 + 1];
    const double du2 = (double)UV[2 * i2] - UV[2 * i0], dv2 = (double)UV[2 * i2 + 1] - UV[2 * i0 + 1];
    const double det = du1 * dv2 - du2 * dv1;
    if (std::fabs(det) < 1e-20) continue;
    const double r = 1.0 / det;
    for (int k = 0; k < 3; ++k) {
      const double tk = (e1[k] * dv2 - e2[k] * dv1) * r;
      const double bk = (e2[k] * du1 - e1[k] * du2) * r;
      for (uint32_t v : {i0, i1, i2}) {
        t[3 * v + k] += tk;
        b[3 * v + k] += bk;
      }
    }
  }
  T->assign(nv * 4, 0.0f);
  for (size_t v = 0; v < nv; ++v) {
    const double n[3] = {N[3 * v], N[3 * v + 1], N[3 * v + 2]};
    double tv[3] = {t[3 * v], t[3 * v + 1], t[3 * v + 2]};
    const double nd = n[0] * tv[0] + n[1] * tv[1] + n[2] * tv[2];
    for (int k = 0; k < 3; ++k) tv[k] -= n[k] * nd;
    double len = std::sqrt(tv[0] * tv[0] + tv[1] * tv[1] + tv[2] * tv[2]);
    float w = 1.0f;
    if (len > 1e-12) {
      const double cx = n[1] * tv[2] - n[2] * tv[1], cy = n[2] * tv[0] - n[0] * tv[2],
                   cz = n[0] * tv[1] - n[1] * tv[0];
      w = (cx * b[3 * v] + cy * b[3 * v + 1] + cz * b[3 * v + 2]) < 0.0 ? -1.0f : 1.0f;
    } else {
      // src/loader.cpp:360-374 fallback frame
      const double a[3] = {std::fabs(n[0]) < 0.9 ? 1.0 : 0.0, std::fabs(n[0]) < 0.9 ? 0.0 : 1.0, 0.0};
      const double d = n[0] * a[0] + n[1] * a[1] + n[2] * a[2];
      for (int k = 0; k < 3; ++k) tv[k] = a[k] - n[k] * d;
      len = std::sqrt(tv[0] * tv[0] + tv[1] * tv[1] + tv[2] * tv[2]);
    }
    for (int k = 0; k < 3; ++k) (*T)[4 * v + k] = len > 0 ? (float)(tv[k] / len) : 0.0f;
    (*T)[4 * v + 3] = w;
  }
}

// src/loader.cpp:144-411 ProcessPrimitive
bool process_primitive(Ctx& c, const Value& prim, const M34& xf) {
  shb_gltf& o = *c.out;
  const int mat = shb::json::integer(prim, "material", -1);
  const Value* attrs = prim.get("attributes");
  if (mat < 0) {
    c.err = "primitive missing material";
    return false;
  }
  if ((size_t)mat >= o.mats.size()) {
    c.err = "primitive material out of range";
    return false;
  }
  auto attr = [&](const char* k) { return attrs ? shb::json::integer(*attrs, k, -1) : -1; };
  if (attr("POSITION") < 0) {
    c.err = "primitive missing POSITION attribute";
    return false;
  }
  if (attr("NORMAL") < 0) {
    c.err = "primitive missing NORMAL attribute (baking requires vertex normals)";
    return false;
  }
  if (attr("TEXCOORD_0") < 0) {
    c.err = "primitive missing TEXCOORD_0 attribute (baking requires UVs)";
    return false;
  }
  std::vector<double> P, N, UV0, UV1, TG, IX;
  size_t nv = 0, n2 = 0;
  if (!read_accessor(c, attr("POSITION"), 3, &P, &nv)) return false;
  if (!read_accessor(c, attr("NORMAL"), 3, &N, &n2) || n2 != nv) {
    if (c.err.empty()) c.err = "NORMAL count differs from POSITION";
    return false;
  }
  if (!read_accessor(c, attr("TEXCOORD_0"), 2, &UV0, &n2) || n2 != nv) {
    if (c.err.empty()) c.err = "TEXCOORD_0 count differs from POSITION";
    return false;
  }
  const bool has_uv1 = attr("TEXCOORD_1") >= 0;
  if (has_uv1 && (!read_accessor(c, attr("TEXCOORD_1"), 2, &UV1, &n2) || n2 != nv)) {
    if (c.err.empty()) c.err = "TEXCOORD_1 count differs from POSITION";
    return false;
  }
  const bool has_tan = attr("TANGENT") >= 0;
  if (has_tan && (!read_accessor(c, attr("TANGENT"), 4, &TG, &n2) || n2 != nv)) {
    if (c.err.empty()) c.err = "TANGENT count differs from POSITION";
    return false;
  }
  std::vector<uint32_t> I;
  const int ia = shb::json::integer(prim, "indices", -1);
  if (ia >= 0) {
    size_t ni = 0;
    if (!read_accessor(c, ia, 1, &IX, &ni, true)) return false;
    for (size_t k = 0; k + 2 < ni; k += 3) {
      const uint32_t a = (uint32_t)IX[k], b = (uint32_t)IX[k + 1], d = (uint32_t)IX[k + 2];
      if (a == b || a == d || b == d) continue;  // loader.cpp:321-353
      if (a >= nv || b >= nv || d >= nv) {
        c.err = "index out of range";
        return false;
      }
      I.push_back(a);
      I.push_back(b);
      I.push_back(d);
    }
  }
  std::vector<float> p(P.begin(), P.end()), n(N.begin(), N.end()), uv(UV0.begin(), UV0.end());
  std::vector<float> tg;
  if (has_tan) tg.assign(TG.begin(), TG.end());
  else generate_tangents(p, n, uv, I, nv, &tg);
  shb_geometry g{};
  g.vertex_offset = (uint32_t)(o.pos.size() / 3);
  g.vertex_count = (uint32_t)nv;
  g.index_offset = (uint32_t)o.idx.size();
  g.index_count = (uint32_t)I.size();
  g.material_id = mat;
  g.flags = SHB_GEOM_HAS_NORMALS | SHB_GEOM_HAS_TEXTURE_UVS | SHB_GEOM_HAS_TANGENTS |
            (has_uv1 ? SHB_GEOM_HAS_LIGHTMAP_UVS : 0u);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j) g.transform[4 * i + j] = (float)xf.m[i][j];
  o.pos.insert(o.pos.end(), p.begin(), p.end());
  o.nrm.insert(o.nrm.end(), n.begin(), n.end());
  o.tuv.insert(o.tuv.end(), uv.begin(), uv.end());
  if (has_uv1) o.luv.insert(o.luv.end(), UV1.begin(), UV1.end());
  else o.luv.insert(o.luv.end(), nv * 2, 0.0f);
  o.tan.insert(o.tan.end(), tg.begin(), tg.end());
  o.idx.insert(o.idx.end(), I.begin(), I.end());
  o.geos.push_back(g);
  if (o.mats[mat].emission_intensity > 0.0f) {  // loader.cpp:391-411
    shb_light l{};
    l.type = SHB_LIGHT_AREA;
    l.direction[2] = -1.0f;
    l.color[0] = l.color[1] = l.color[2] = 1.0f;
    l.intensity = o.mats[mat].emission_intensity;
    l.cos_inner_cone = 1.0f;
    l.cos_outer_cone = 0.70710678118654752440f;
    l.geometry_index = (int32_t)o.geos.size() - 1;
    l.material_index = mat;
    o.lights.push_back(l);
  }
  return true;
}

// src/loader.cpp:545-610 ProcessLight
void process_light(Ctx& c, const Value& lo, const M34& xf) {
  if (!lo.is_object()) return;
  shb_light l{};
  l.type = SHB_LIGHT_POINT;
  for (int i = 0; i < 3; ++i) l.position[i] = (float)xf.m[i][3];
  double R[3][3];
  polar_rotation(xf, R);
  float d[3] = {(float)-R[0][2], (float)-R[1][2], (float)-R[2][2]};
  const float len2 = d[0] * d[0] + (d[1] * d[1] + d[2] * d[2]);
  if (len2 > 1e-8f && std::isfinite(len2)) {
    const float len = std::sqrt(len2);
    for (int i = 0; i < 3; ++i) d[i] /= len;
  } else {
    d[0] = d[1] = 0.0f;
    d[2] = -1.0f;
  }
  std::memcpy(l.direction, d, sizeof(d));
  const std::string type = shb::json::string(lo, "type");
  if (type == "directional") l.type = SHB_LIGHT_DIRECTIONAL;
  else if (type == "spot") l.type = SHB_LIGHT_SPOT;
  l.color[0] = l.color[1] = l.color[2] = 1.0f;
  const Value* col = lo.get("color");
  if (col && col->is_array() && col->size() == 3)
    for (int i = 0; i < 3; ++i) l.color[i] = (float)(*col)[i].number_or(1);
  l.intensity = (float)shb::json::num(lo, "intensity", 1.0);
  l.cos_inner_cone = 1.0f;
  l.cos_outer_cone = 0.70710678118654752440f;
  const Value* spot = lo.get("spot");
  if (l.type == SHB_LIGHT_SPOT && spot) {
    if (spot->get("innerConeAngle"))
      l.cos_inner_cone = std::cos((float)spot->get("innerConeAngle")->number_or(0));
    if (spot->get("outerConeAngle"))
      l.cos_outer_cone = std::cos((float)spot->get("outerConeAngle")->number_or(0.785398));
  }
  l.geometry_index = -1;
  l.material_index = -1;
  c.out->lights.push_back(l);
}

// src/loader.cpp:612-654 TraverseNodes
bool traverse(Ctx& c, int ni, const M34& parent, int depth) {
  const Value* node = arr_at(c.root, "nodes", ni);
  if (!node) {
    c.err = "node " + std::to_string(ni) + " out of range";
    return false;
  }
  if (depth > 1000) {
    c.err = "node hierarchy too deep (cycle?)";
    return false;
  }
  const M34 xf = parent * node_transform(*node);
  const int mesh = shb::json::integer(*node, "mesh", -1);
  if (mesh >= 0) {
    const Value* m = arr_at(c.root, "meshes", mesh);
    const Value* prims = m ? m->get("primitives") : nullptr;
    if (!m) {
      c.err = "mesh out of range";
      return false;
    }
    if (prims)
      for (size_t k = 0; k < prims->size(); ++k)
        if (!process_primitive(c, (*prims)[k], xf)) return false;
  }
  const Value* ext = node->get("extensions");
  const Value* lp = ext ? ext->get("KHR_lights_punctual") : nullptr;
  if (lp && lp->get("light")) {
    const int li = (int)lp->get("light")->number_or(-1);
    const Value* rext = c.root->get("extensions");
    const Value* rlp = rext ? rext->get("KHR_lights_punctual") : nullptr;
    const Value* arr = rlp ? rlp->get("lights") : nullptr;
    if (arr && arr->is_array() && li >= 0 && (size_t)li < arr->size())
      process_light(c, (*arr)[li], xf);
  }
  const Value* ch = node->get("children");
  if (ch && ch->is_array())
    for (size_t k = 0; k < ch->size(); ++k)
      if (!traverse(c, (int)(*ch)[k].number_or(-1), xf, depth + 1)) return false;
  return true;
}

// src/loader.cpp:656-789 ProcessEnvironment
bool process_environment(Ctx& c) {
  shb_gltf& o = *c.out;
  const Value* extras = c.root->get("extras");
  const Value* sky = extras ? extras->get("skybox") : nullptr;
  if (sky && sky->is_string()) {
    std::vector<uint8_t> bytes;
    std::string path;
    if (load_uri(c, sky->str, &bytes, &path)) {
      std::string e;
      if (shb::is_hdr(bytes.data(), bytes.size())) {
        shb::ImageF im;
        if (shb::decode_hdr(bytes.data(), bytes.size(), &im, &e)) {
          o.env_type = SHB_ENV_TEXTURE;
          o.env_w = im.width;
          o.env_h = im.height;
          o.env_rgb = std::move(im.rgb);
          o.env_path = path;
          return true;
        }
      } else {
        shb::Image8 im;
        if (shb::decode_image8(bytes.data(), bytes.size(), &im, &e)) {
          o.env_type = SHB_ENV_TEXTURE;
          o.env_w = im.width;
          o.env_h = im.height;
          o.env_rgb.resize((size_t)im.width * im.height * 3);
          for (size_t i = 0; i < (size_t)im.width * im.height; ++i)
            for (int k = 0; k < 3; ++k) o.env_rgb[i * 3 + k] = im.rgba[i * 4 + k] / 255.0f * 1.0f;
          o.env_path = path;
          return true;
        }
      }
    }
    c.err.clear();  // a skybox that cannot be read falls through to the sun
  }
  int best = -1;
  float best_i = -1.0f;
  for (size_t i = 0; i < o.lights.size(); ++i)
    if (o.lights[i].type == SHB_LIGHT_DIRECTIONAL && o.lights[i].intensity > best_i) {
      best_i = o.lights[i].intensity;
      best = (int)i;
    }
  if (best >= 0) {
    o.env_type = SHB_ENV_PREETHAM;
    for (int k = 0; k < 3; ++k) o.sun[k] = -o.lights[best].direction[k];
    o.sun_intensity = o.lights[best].intensity;
  }
  return true;
}

bool parse_glb(Ctx& c, const std::vector<uint8_t>& f, std::string* json) {
  auto u32 = [&](size_t p) {
    uint32_t v;
    std::memcpy(&v, &f[p], 4);
    return v;
  };
  if (f.size() < 20 || u32(0) != 0x46546C67u) {
    c.err = "not a GLB file";
    return false;
  }
  if (u32(4) != 2) {
    c.err = "GLB version is not 2";
    return false;
  }
  const size_t total = std::min<size_t>(u32(8), f.size());
  size_t p = 12;
  bool have_json = false;
  while (p + 8 <= total) {
    const uint32_t len = u32(p), type = u32(p + 4);
    if (p + 8 + (size_t)len > total) {
      c.err = "GLB chunk past the end of the file";
      return false;
    }
    if (type == 0x4E4F534Au && !have_json) {
      json->assign(reinterpret_cast<const char*>(&f[p + 8]), len);
      have_json = true;
    } else if (type == 0x004E4942u && !c.has_glb_bin) {
      c.glb_bin.assign(f.begin() + p + 8, f.begin() + p + 8 + len);
      c.has_glb_bin = true;
    }
    p += 8 + (size_t)len;
  }
  if (!have_json) c.err = "GLB without a JSON chunk";
  return have_json;
}

bool load(const std::string& path, shb_gltf* out, std::string* err) {
  Ctx c;
  c.out = out;
  const size_t slash = path.find_last_of('/');
  c.dir = slash == std::string::npos ? "" : path.substr(0, slash + 1);
  std::vector<uint8_t> file;
  if (!read_file(path, &file)) {
    *err = "cannot read " + path;
    return false;
  }
  std::string text;
  const bool glb = path.size() >= 4 && path.compare(path.size() - 4, 4, ".glb") == 0;
  if (glb) {
    if (!parse_glb(c, file, &text)) {
      *err = c.err;
      return false;
    }
  } else {
    text.assign(file.begin(), file.end());
  }
  Value root;
  if (!shb::json::parse(text, &root, err)) return false;
  if (!root.is_object()) {
    *err = "glTF root is not an object";
    return false;
  }
  c.root = &root;
  const Value* asset = root.get("asset");
  if (!asset || shb::json::string(*asset, "version").rfind("2", 0) != 0) {
    *err = "not a glTF 2.x asset";
    return false;
  }
  if (!load_buffers(c) || !process_materials(c)) {
    *err = c.err;
    return false;
  }
  const Value* scenes = root.get("scenes");
  const int si = shb::json::integer(root, "scene", -1);
  const Value* sc = arr_at(&root, "scenes", si > -1 ? si : 0);
  if (!scenes || !sc) {
    *err = "glTF has no scene";
    return false;
  }
  const Value* nodes = sc->get("nodes");
  if (nodes && nodes->is_array())
    for (size_t k = 0; k < nodes->size(); ++k)
      if (!traverse(c, (int)(*nodes)[k].number_or(-1), M34(), 0)) {
        *err = "failed to process the scene graph: " + c.err;
        return false;
      }
  process_environment(c);
  shb_scene_desc& d = out->desc;
  d.num_vertices = (uint32_t)(out->pos.size() / 3);
  d.positions = out->pos.data();
  d.normals = out->nrm.data();
  d.texture_uvs = out->tuv.data();
  d.lightmap_uvs = out->luv.data();
  d.tangents = out->tan.data();
  d.num_indices = (uint32_t)out->idx.size();
  d.indices = out->idx.data();
  d.num_geometries = (uint32_t)out->geos.size();
  d.geometries = out->geos.data();
  d.num_materials = (uint32_t)out->mats.size();
  d.materials = out->mats.data();
  d.texel_pool_size = out->texels.size();
  d.texel_pool = out->texels.data();
  d.num_lights = (uint32_t)out->lights.size();
  d.lights = out->lights.data();
  return true;
}

void put_err(const std::string& e, char* err, size_t len) {
  if (!err || !len) return;
  std::strncpy(err, e.c_str(), len - 1);
  err[len - 1] = 0;
}

}  // namespace

extern "C" {

int shb_load_gltf(const char* path, shb_gltf** out, char* err, size_t err_len) {
  if (!path || !out) {
    put_err("bad arguments", err, err_len);
    return SHB_ERR_INVALID_ARG;
  }
  *out = nullptr;
  auto* g = new (std::nothrow) shb_gltf();
  if (!g) return SHB_ERR_OOM;
  std::string e;
  bool ok = false;
  try {
    ok = load(path, g, &e);
  } catch (const std::bad_alloc&) {
    e = "out of host memory";
  }
  if (!ok) {
    delete g;
    put_err(e, err, err_len);
    return SHB_ERR_INVALID_ARG;
  }
  *out = g;
  return SHB_OK;
}

const shb_scene_desc* shb_gltf_scene(const shb_gltf* g) { return g ? &g->desc : nullptr; }

int shb_gltf_environment(const shb_gltf* g, shb_environment_desc* e) {
  if (!g || !e) return SHB_ERR_INVALID_ARG;
  std::memset(e, 0, sizeof(*e));
  e->type = g->env_type;
  e->width = g->env_w;
  e->height = g->env_h;
  e->channels = g->env_type == SHB_ENV_TEXTURE ? 3 : 0;
  e->pixels = g->env_rgb.empty() ? nullptr : g->env_rgb.data();
  std::memcpy(e->sun_direction, g->sun, sizeof(g->sun));
  e->sun_intensity = g->sun_intensity;
  e->path = g->env_path.empty() ? nullptr : g->env_path.c_str();
  return SHB_OK;
}

const char* shb_gltf_material_name(const shb_gltf* g, uint32_t i) {
  return g && i < g->mat_names.size() ? g->mat_names[i].c_str() : nullptr;
}

const char* shb_gltf_albedo_path(const shb_gltf* g, uint32_t i) {
  return g && i < g->albedo_paths.size() && !g->albedo_paths[i].empty()
             ? g->albedo_paths[i].c_str()
             : nullptr;
}

void shb_gltf_free(shb_gltf* g) { delete g; }

}  // extern "C"
