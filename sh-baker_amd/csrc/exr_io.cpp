// OpenEXR output of a baked SH lightmap (host side), the three layouts of
// the reference's SaveSHLightMap (src/saver.cpp:103-422): one 28-channel
// file, nine per-coefficient files (L0 carries visibility in A), or three
// RGBA files of packed luminance. Channels are stored as HALF with ZIP
// compression, as the reference requests; the channel list is written in
// name order (the file format requires it; readers look channels up by name,
// so the reference's unsorted request stores the same data).
//
// The writer follows the OpenEXR file layout specification: magic and
// version, a header of typed attributes (channel list sorted by name,
// compression, windows, line order, aspect, screen window), a table of
// chunk offsets, then blocks of 16 scanlines; each block holds, per
// scanline, each channel's row of 2-byte halves, split into even and odd
// bytes, delta-coded byte to byte (+128, mod 256) and deflated with zlib
// (a block that does not shrink is stored raw).
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "shb_ctx.h"

namespace {

// IEEE binary32 -> binary16, round to nearest, ties to even; overflow to
// infinity, NaN kept quiet, subnormal halves produced exactly.
uint16_t float_to_half(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t exp = (x >> 23) & 0xffu;
  uint32_t mant = x & 0x7fffffu;
  if (exp == 0xffu) return (uint16_t)(sign | 0x7c00u | (mant ? 0x200u : 0u));
  int e = (int)exp - 127 + 15;
  if (e >= 31) return (uint16_t)(sign | 0x7c00u);
  if (e <= 0) {
    if (e < -10) return (uint16_t)sign;  // rounds to zero (below half of the smallest subnormal)
    mant |= 0x800000u;                   // implicit bit
    const int shift = 14 - e;            // 24-bit significand -> subnormal half
    uint32_t h = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u);
    const uint32_t half = 1u << (shift - 1);
    if (rem > half || (rem == half && (h & 1u))) ++h;
    return (uint16_t)(sign | h);
  }
  uint32_t h = ((uint32_t)e << 10) | (mant >> 13);
  const uint32_t rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;  // may carry into the exponent
  return (uint16_t)(sign | h);
}

void put_u32(std::vector<uint8_t>& b, uint32_t v) {
  for (int i = 0; i < 4; ++i) b.push_back((uint8_t)(v >> (8 * i)));
}
void put_u64(std::vector<uint8_t>& b, uint64_t v) {
  for (int i = 0; i < 8; ++i) b.push_back((uint8_t)(v >> (8 * i)));
}
void put_f32(std::vector<uint8_t>& b, float f) {
  uint32_t v;
  std::memcpy(&v, &f, 4);
  put_u32(b, v);
}
void put_str(std::vector<uint8_t>& b, const std::string& s) {
  b.insert(b.end(), s.begin(), s.end());
  b.push_back(0);
}
void attr(std::vector<uint8_t>& b, const char* name, const char* type,
          const std::vector<uint8_t>& value) {
  put_str(b, name);
  put_str(b, type);
  put_u32(b, (uint32_t)value.size());
  b.insert(b.end(), value.begin(), value.end());
}

// planes[c] holds channel c's width*height floats; names[c] its name.
bool write_exr(const std::string& path, int w, int h, const std::vector<std::string>& names,
               const std::vector<const float*>& planes, std::string* err) {
  const int nc = (int)names.size();
  std::vector<int> order(nc);
  for (int i = 0; i < nc; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return names[a] < names[b]; });
  std::vector<uint8_t> f;
  put_u32(f, 20000630u);  // magic
  put_u32(f, 2u);         // version 2, single-part scanline
  {
    std::vector<uint8_t> v;
    for (int k : order) {
      put_str(v, names[k]);
      put_u32(v, 1u);  // HALF
      v.push_back(0);  // pLinear
      v.push_back(0);
      v.push_back(0);
      v.push_back(0);  // reserved
      put_u32(v, 1u);  // x sampling
      put_u32(v, 1u);  // y sampling
    }
    v.push_back(0);
    attr(f, "channels", "chlist", v);
  }
  attr(f, "compression", "compression", {3});  // ZIP, 16 scanlines per block
  {
    std::vector<uint8_t> v;
    put_u32(v, 0);
    put_u32(v, 0);
    put_u32(v, (uint32_t)(w - 1));
    put_u32(v, (uint32_t)(h - 1));
    attr(f, "dataWindow", "box2i", v);
    attr(f, "displayWindow", "box2i", v);
  }
  attr(f, "lineOrder", "lineOrder", {0});
  {
    std::vector<uint8_t> v;
    put_f32(v, 1.f);
    attr(f, "pixelAspectRatio", "float", v);
  }
  {
    std::vector<uint8_t> v;
    put_f32(v, 0.f);
    put_f32(v, 0.f);
    attr(f, "screenWindowCenter", "v2f", v);
  }
  {
    std::vector<uint8_t> v;
    put_f32(v, 1.f);
    attr(f, "screenWindowWidth", "float", v);
  }
  f.push_back(0);  // end of header
  const int lines = 16;
  const int nblocks = (h + lines - 1) / lines;
  const size_t table = f.size();
  for (int i = 0; i < nblocks; ++i) put_u64(f, 0);
  std::vector<uint8_t> raw, split, packed;
  for (int bi = 0; bi < nblocks; ++bi) {
    const int y0 = bi * lines, y1 = std::min(h, y0 + lines);
    raw.clear();
    for (int y = y0; y < y1; ++y)
      for (int k : order) {
        const float* row = planes[k] + (size_t)y * w;
        for (int x = 0; x < w; ++x) {
          uint16_t hv = float_to_half(row[x]);
          raw.push_back((uint8_t)(hv & 0xff));
          raw.push_back((uint8_t)(hv >> 8));
        }
      }
    const size_t n = raw.size();
    split.assign(n, 0);
    size_t e = 0, o = (n + 1) / 2;
    for (size_t i = 0; i < n; ++i) split[(i & 1) ? o++ : e++] = raw[i];
    for (size_t i = n; i-- > 1;) split[i] = (uint8_t)(split[i] - split[i - 1] + 128);
    uLongf clen = compressBound((uLong)n);
    packed.resize(clen);
    if (compress2(packed.data(), &clen, split.data(), (uLong)n, Z_DEFAULT_COMPRESSION) != Z_OK) {
      *err = "zlib compression failed";
      return false;
    }
    const uint64_t at = f.size();
    std::memcpy(&f[table + 8 * (size_t)bi], &at, 8);  // little-endian host
    put_u32(f, (uint32_t)y0);
    if (clen < n) {
      put_u32(f, (uint32_t)clen);
      f.insert(f.end(), packed.begin(), packed.begin() + clen);
    } else {
      put_u32(f, (uint32_t)n);
      f.insert(f.end(), raw.begin(), raw.end());
    }
  }
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) {
    *err = "cannot open " + path + " for writing";
    return false;
  }
  const bool wrote = std::fwrite(f.data(), 1, f.size(), fp) == f.size();
  const bool ok = (std::fclose(fp) == 0) && wrote;
  if (!ok) *err = "short write to " + path;
  return ok;
}

// path "dir/name.ext" -> "dir/name" + suffix + ".ext" (saver.cpp:210-221)
std::string with_suffix(const std::string& path, const std::string& suffix) {
  size_t slash = path.find_last_of('/');
  size_t dot = path.find_last_of('.');
  if (dot == std::string::npos || (slash != std::string::npos && dot < slash))
    return path + suffix;
  return path.substr(0, dot) + suffix + path.substr(dot);
}

const char* kCoeffNames[9] = {"L0", "L1m1", "L10", "L11", "L2m2", "L2m1", "L20", "L21", "L22"};

}  // namespace

extern "C" {

int shb_save_lightmap(const char* path, int32_t width, int32_t height, const float* sh,
                      const float* vis, int32_t mode, char* err_out, size_t err_len) {
  std::string err;
  auto done = [&](int rc) {
    if (rc != SHB_OK && err_out && err_len) {
      std::strncpy(err_out, err.c_str(), err_len - 1);
      err_out[err_len - 1] = 0;
    }
    return rc;
  };
  if (!path || !sh || width <= 0 || height <= 0) {
    err = "bad arguments";
    return done(SHB_ERR_INVALID_ARG);
  }
  if (mode < SHB_SAVE_COMBINED || mode > SHB_SAVE_LUMINANCE_PACKED) {
    err = "unknown save mode";
    return done(SHB_ERR_INVALID_ARG);
  }
  const size_t n = (size_t)width * height;
  // channel planes: the 27 SH values and visibility (zeros when absent)
  std::vector<std::vector<float>> c27(27, std::vector<float>(n));
  for (size_t p = 0; p < n; ++p)
    for (int k = 0; k < 27; ++k) c27[k][p] = sh[27 * p + k];
  std::vector<float> v(n, 0.f);
  if (vis) std::copy(vis, vis + n, v.begin());
  const std::string P(path);
  if (mode == SHB_SAVE_COMBINED) {  // saver.cpp:108-199
    std::vector<std::string> names;
    std::vector<const float*> planes;
    const char* rgb[3] = {".R", ".G", ".B"};
    for (int i = 0; i < 9; ++i)
      for (int c = 0; c < 3; ++c) {
        names.push_back(std::string(kCoeffNames[i]) + rgb[c]);
        planes.push_back(c27[3 * i + c].data());
      }
    names.push_back("EnvVisibility");
    planes.push_back(v.data());
    return done(write_exr(P, width, height, names, planes, &err) ? SHB_OK : SHB_ERR_DEVICE);
  }
  if (mode == SHB_SAVE_SPLIT) {  // saver.cpp:201-296
    for (int i = 0; i < 9; ++i) {
      std::vector<std::string> names = {"R", "G", "B"};
      std::vector<const float*> planes = {c27[3 * i].data(), c27[3 * i + 1].data(),
                                          c27[3 * i + 2].data()};
      if (i == 0) {
        names.push_back("A");
        planes.push_back(v.data());
      }
      if (!write_exr(with_suffix(P, std::string("_") + kCoeffNames[i]), width, height, names,
                     planes, &err))
        return done(SHB_ERR_DEVICE);
    }
    return done(SHB_OK);
  }
  // packed luminance (saver.cpp:298-401): file 0 = L0 RGB + visibility,
  // file 1 = luma of L1m1 L10 L11 L2m2, file 2 = luma of L2m1 L20 L21 L22;
  // luma = 0.2126 R + 0.7152 G + 0.0722 B summed as R*wr + (G*wg + B*wb)
  const float wr = 0.2126f, wg = 0.7152f, wb = 0.0722f;
  auto luma = [&](int coeff, std::vector<float>& out) {
    out.resize(n);
    for (size_t p = 0; p < n; ++p)
      out[p] = c27[3 * coeff][p] * wr + (c27[3 * coeff + 1][p] * wg + c27[3 * coeff + 2][p] * wb);
  };
  std::vector<std::vector<float>> lum(8);
  for (int k = 0; k < 8; ++k) luma(1 + k, lum[k]);
  const std::vector<std::string> names = {"R", "G", "B", "A"};
  for (int fi = 0; fi < 3; ++fi) {
    std::vector<const float*> planes;
    if (fi == 0)
      planes = {c27[0].data(), c27[1].data(), c27[2].data(), v.data()};
    else
      for (int c = 0; c < 4; ++c) planes.push_back(lum[4 * (fi - 1) + c].data());
    if (!write_exr(with_suffix(P, "_packed_" + std::to_string(fi)), width, height, names, planes,
                   &err))
      return done(SHB_ERR_DEVICE);
  }
  return done(SHB_OK);
}

}  // extern "C"
