// Host image decoders for the glTF loader (see image_decode.h).
//
// PNG follows the PNG specification (ISO/IEC 15948): zlib-inflated IDAT,
// the five scanline filters, Adam7, tRNS colour keys and palette alpha;
// 16-bit samples keep their high byte, low bit depths are scaled to 8 bits.
// JPEG follows ITU-T T.81 for Huffman sequential DCT (SOF0/SOF1, 8-bit):
// canonical Huffman tables, DC prediction, restart intervals, a separable
// float IDCT, triangle-filter ("fancy") upsampling of 2:1 chroma, and the
// JFIF YCbCr -> RGB transform. Progressive and arithmetic-coded files are
// rejected. Radiance RGBE follows the format of Ward's Radiance: a text
// header, "-Y h +X w", adaptive run-length scanlines, value = m * 2^(e-136).
#include "image_decode.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstring>

namespace shb {

namespace {

bool fail(std::string* err, const std::string& m) {
  if (err) *err = m;
  return false;
}

uint32_t be32(const uint8_t* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

// ------------------------------------------------------------------ PNG
uint8_t paeth(int a, int b, int c) {
  int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  if (pa <= pb && pa <= pc) return (uint8_t)a;
  if (pb <= pc) return (uint8_t)b;
  return (uint8_t)c;
}

}  // namespace

bool decode_png(const uint8_t* d, size_t n, Image8* out, std::string* err) {
  static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (n < 8 || std::memcmp(d, sig, 8) != 0) return fail(err, "PNG: bad signature");
  size_t p = 8;
  uint32_t w = 0, h = 0;
  int depth = 0, ctype = -1, interlace = 0;
  std::vector<uint8_t> idat, plte, trns;
  bool have_ihdr = false;
  while (p + 12 <= n) {
    const uint32_t len = be32(d + p);
    const uint8_t* type = d + p + 4;
    const uint8_t* body = d + p + 8;
    if ((size_t)len > n - p - 12) return fail(err, "PNG: truncated chunk");
    if (!std::memcmp(type, "IHDR", 4)) {
      if (len < 13) return fail(err, "PNG: bad IHDR");
      w = be32(body);
      h = be32(body + 4);
      depth = body[8];
      ctype = body[9];
      if (body[10] != 0 || body[11] != 0) return fail(err, "PNG: unknown compression/filter");
      interlace = body[12];
      have_ihdr = true;
    } else if (!std::memcmp(type, "PLTE", 4)) {
      plte.assign(body, body + len);
    } else if (!std::memcmp(type, "tRNS", 4)) {
      trns.assign(body, body + len);
    } else if (!std::memcmp(type, "IDAT", 4)) {
      idat.insert(idat.end(), body, body + len);
    } else if (!std::memcmp(type, "IEND", 4)) {
      break;
    }
    p += 12 + (size_t)len;
  }
  if (!have_ihdr || w == 0 || h == 0) return fail(err, "PNG: missing IHDR");
  if (w > (1u << 24) || h > (1u << 24) || (uint64_t)w * h > (1ull << 28))
    return fail(err, "PNG: image too large");
  int channels;
  switch (ctype) {
    case 0: channels = 1; break;
    case 2: channels = 3; break;
    case 3: channels = 1; break;
    case 4: channels = 2; break;
    case 6: channels = 4; break;
    default: return fail(err, "PNG: bad colour type");
  }
  const bool ok_depth =
      (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
      (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
      ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
  if (!ok_depth) return fail(err, "PNG: bad bit depth for colour type");
  if (ctype == 3 && plte.size() < 3) return fail(err, "PNG: palette image without PLTE");
  if (interlace > 1) return fail(err, "PNG: unknown interlace method");
  const int bits = channels * depth;
  const int bpp = std::max(1, bits / 8);
  static const int ax0[7] = {0, 4, 0, 2, 0, 1, 0}, ay0[7] = {0, 0, 4, 0, 2, 0, 1};
  static const int adx[7] = {8, 8, 4, 4, 2, 2, 1}, ady[7] = {8, 8, 8, 4, 4, 2, 2};
  const int npass = interlace ? 7 : 1;
  uint32_t pw[7], ph[7];
  uint64_t total = 0;
  for (int k = 0; k < npass; ++k) {
    const int x0 = interlace ? ax0[k] : 0, y0 = interlace ? ay0[k] : 0;
    const int dx = interlace ? adx[k] : 1, dy = interlace ? ady[k] : 1;
    pw[k] = w > (uint32_t)x0 ? (w - x0 + dx - 1) / dx : 0;
    ph[k] = h > (uint32_t)y0 ? (h - y0 + dy - 1) / dy : 0;
    if (pw[k] && ph[k]) total += (uint64_t)ph[k] * (1 + ((uint64_t)pw[k] * bits + 7) / 8);
  }
  std::vector<uint8_t> raw((size_t)total);
  uLongf rl = (uLongf)total;
  if (uncompress(raw.data(), &rl, idat.data(), (uLong)idat.size()) != Z_OK || rl != total)
    return fail(err, "PNG: bad image data");
  out->width = (int)w;
  out->height = (int)h;
  out->rgba.assign((size_t)w * h * 4, 0);
  // colour key (tRNS for grey / RGB), compared on raw samples
  const bool key_grey = ctype == 0 && trns.size() >= 2;
  const bool key_rgb = ctype == 2 && trns.size() >= 6;
  const unsigned kg = key_grey ? (trns[0] << 8 | trns[1]) : 0;
  const unsigned kr = key_rgb ? (trns[0] << 8 | trns[1]) : 0;
  const unsigned kgn = key_rgb ? (trns[2] << 8 | trns[3]) : 0;
  const unsigned kb = key_rgb ? (trns[4] << 8 | trns[5]) : 0;
  static const int scale[17] = {0, 0xff, 0x55, 0, 0x11, 0, 0, 0, 0x01, 0, 0, 0, 0, 0, 0, 0, 0};
  size_t off = 0;
  std::vector<uint8_t> prev;
  for (int k = 0; k < npass; ++k) {
    if (!pw[k] || !ph[k]) continue;
    const int x0 = interlace ? ax0[k] : 0, y0 = interlace ? ay0[k] : 0;
    const int dx = interlace ? adx[k] : 1, dy = interlace ? ady[k] : 1;
    const size_t rowbytes = ((size_t)pw[k] * bits + 7) / 8;
    prev.assign(rowbytes, 0);
    for (uint32_t j = 0; j < ph[k]; ++j) {
      const int ft = raw[off];
      uint8_t* row = &raw[off + 1];
      off += 1 + rowbytes;
      for (size_t i = 0; i < rowbytes; ++i) {
        const int a = i >= (size_t)bpp ? row[i - bpp] : 0;
        const int b = prev[i];
        const int c = i >= (size_t)bpp ? prev[i - bpp] : 0;
        switch (ft) {
          case 0: break;
          case 1: row[i] = (uint8_t)(row[i] + a); break;
          case 2: row[i] = (uint8_t)(row[i] + b); break;
          case 3: row[i] = (uint8_t)(row[i] + ((a + b) >> 1)); break;
          case 4: row[i] = (uint8_t)(row[i] + paeth(a, b, c)); break;
          default: return fail(err, "PNG: bad filter type");
        }
      }
      std::memcpy(prev.data(), row, rowbytes);
      const uint32_t yy = y0 + j * dy;
      for (uint32_t i = 0; i < pw[k]; ++i) {
        const uint32_t xx = x0 + i * dx;
        uint8_t* o = &out->rgba[((size_t)yy * w + xx) * 4];
        unsigned s[4] = {0, 0, 0, 0};  // raw samples (full depth)
        if (depth < 8) {
          const size_t bit = (size_t)i * depth;
          s[0] = (row[bit >> 3] >> (8 - depth - (bit & 7))) & ((1u << depth) - 1);
        } else {
          for (int c = 0; c < channels; ++c)
            s[c] = depth == 8 ? row[(size_t)i * channels + c]
                              : (unsigned)(row[((size_t)i * channels + c) * 2] << 8 |
                                           row[((size_t)i * channels + c) * 2 + 1]);
        }
        auto to8 = [&](unsigned v) -> uint8_t {
          return depth == 16 ? (uint8_t)(v >> 8) : depth == 8 ? (uint8_t)v : (uint8_t)(v * scale[depth]);
        };
        if (ctype == 3) {
          const unsigned idx = s[0];
          if (idx * 3 + 2 >= plte.size()) return fail(err, "PNG: palette index out of range");
          o[0] = plte[idx * 3];
          o[1] = plte[idx * 3 + 1];
          o[2] = plte[idx * 3 + 2];
          o[3] = idx < trns.size() ? trns[idx] : 255;
        } else if (ctype == 0 || ctype == 4) {
          o[0] = o[1] = o[2] = to8(s[0]);
          o[3] = ctype == 4 ? to8(s[1]) : (key_grey && s[0] == kg ? 0 : 255);
        } else {
          o[0] = to8(s[0]);
          o[1] = to8(s[1]);
          o[2] = to8(s[2]);
          o[3] = ctype == 6 ? to8(s[3])
                            : (key_rgb && s[0] == kr && s[1] == kgn && s[2] == kb ? 0 : 255);
        }
      }
    }
  }
  return true;
}

// ----------------------------------------------------------------- JPEG
namespace {

const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                             12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                             35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                             58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huff {
  bool defined = false;
  uint8_t counts[17] = {0};
  uint8_t vals[256] = {0};
  int mincode[17] = {0}, maxcode[17] = {0}, valptr[17] = {0};
  void build() {
    int code = 0, k = 0;
    for (int len = 1; len <= 16; ++len) {
      valptr[len] = k;
      mincode[len] = code;
      code += counts[len];
      k += counts[len];
      maxcode[len] = counts[len] ? code - 1 : -1;
      code <<= 1;
    }
    defined = true;
  }
};

struct Comp {
  int id = 0, h = 1, v = 1, tq = 0, td = 0, ta = 0;
  int bw = 0, bh = 0;  // plane size in blocks
  std::vector<uint8_t> plane;
  int pred = 0;
};

struct Bits {
  const uint8_t* d;
  size_t n, p;
  uint32_t acc = 0;
  int cnt = 0;
  bool marker = false;  // a marker was reached: feed zeros
  void fill() {
    while (cnt <= 24) {
      uint32_t byte = 0;
      if (!marker && p < n) {
        byte = d[p];
        if (byte == 0xFF) {
          const uint8_t nx = p + 1 < n ? d[p + 1] : 0;
          if (nx == 0x00) {
            p += 2;
          } else {
            marker = true;
            byte = 0;
          }
        } else {
          ++p;
        }
      }
      acc |= byte << (24 - cnt);
      cnt += 8;
    }
  }
  int bit() {
    if (cnt < 1) fill();
    const int b = (int)(acc >> 31);
    acc <<= 1;
    --cnt;
    return b;
  }
  int get(int s) {
    if (s == 0) return 0;
    if (cnt < s) fill();
    const int v = (int)(acc >> (32 - s));
    acc <<= s;
    cnt -= s;
    return v;
  }
  void reset() {  // byte-align after a restart marker
    acc = 0;
    cnt = 0;
    marker = false;
  }
};

int extend(int v, int s) { return s && v < (1 << (s - 1)) ? v - (1 << s) + 1 : v; }

bool huff_decode(Bits& b, const Huff& t, int* out) {
  int code = 0;
  for (int len = 1; len <= 16; ++len) {
    code = (code << 1) | b.bit();
    if (t.maxcode[len] >= 0 && code <= t.maxcode[len]) {
      *out = t.vals[t.valptr[len] + code - t.mincode[len]];
      return true;
    }
  }
  return false;
}

// T[x][u] = C(u) / 2 * cos((2x + 1) u pi / 16), C(0) = 1 / sqrt(2)
struct IdctTable {
  float t[8][8];
  IdctTable() {
    for (int x = 0; x < 8; ++x)
      for (int u = 0; u < 8; ++u)
        t[x][u] = (float)((u == 0 ? 1.0 / std::sqrt(2.0) : 1.0) / 2.0 *
                          std::cos((2 * x + 1) * u * 3.14159265358979323846 / 16.0));
  }
};

void idct_block(const int* coef, uint8_t* dst, int stride) {
  static const IdctTable T;
  float tmp[8][8];
  for (int v = 0; v < 8; ++v)
    for (int x = 0; x < 8; ++x) {
      float s = 0.f;
      for (int u = 0; u < 8; ++u) s += T.t[x][u] * (float)coef[v * 8 + u];
      tmp[v][x] = s;
    }
  for (int y = 0; y < 8; ++y)
    for (int x = 0; x < 8; ++x) {
      float s = 0.f;
      for (int v = 0; v < 8; ++v) s += T.t[y][v] * tmp[v][x];
      const int q = (int)std::lround(s + 128.0f);
      dst[y * stride + x] = (uint8_t)std::min(255, std::max(0, q));
    }
}

}  // namespace

bool decode_jpeg(const uint8_t* d, size_t n, Image8* out, std::string* err) {
  if (n < 4 || d[0] != 0xFF || d[1] != 0xD8) return fail(err, "JPEG: bad signature");
  uint16_t q[4][64] = {{0}};
  bool qdef[4] = {false, false, false, false};
  Huff hdc[4], hac[4];
  std::vector<Comp> comps;
  int width = 0, height = 0, restart = 0, hmax = 1, vmax = 1;
  bool frame = false, done = false;
  size_t p = 2;
  while (p + 4 <= n && !done) {
    if (d[p] != 0xFF) return fail(err, "JPEG: expected a marker");
    const uint8_t m = d[p + 1];
    if (m == 0xFF) {  // fill byte
      ++p;
      continue;
    }
    if (m == 0xD8 || (m >= 0xD0 && m <= 0xD7) || m == 0x01) {
      p += 2;
      continue;
    }
    if (m == 0xD9) break;
    const size_t len = (size_t)d[p + 2] << 8 | d[p + 3];
    if (len < 2 || p + 2 + len > n) return fail(err, "JPEG: truncated segment");
    const uint8_t* s = d + p + 4;
    const size_t sl = len - 2;
    if (m == 0xDB) {  // DQT
      size_t k = 0;
      while (k < sl) {
        const int pq = s[k] >> 4, tq = s[k] & 15;
        if (tq > 3) return fail(err, "JPEG: bad quantisation table id");
        ++k;
        for (int i = 0; i < 64; ++i) {
          if (k + (pq ? 2 : 1) > sl) return fail(err, "JPEG: truncated DQT");
          q[tq][i] = pq ? (uint16_t)(s[k] << 8 | s[k + 1]) : s[k];
          k += pq ? 2 : 1;
        }
        qdef[tq] = true;
      }
    } else if (m == 0xC4) {  // DHT
      size_t k = 0;
      while (k < sl) {
        const int tc = s[k] >> 4, th = s[k] & 15;
        if (tc > 1 || th > 3 || k + 17 > sl) return fail(err, "JPEG: bad Huffman table");
        Huff& t = tc ? hac[th] : hdc[th];
        int total = 0;
        for (int i = 1; i <= 16; ++i) {
          t.counts[i] = s[k + i];
          total += t.counts[i];
        }
        k += 17;
        if (total > 256 || k + total > sl) return fail(err, "JPEG: bad Huffman table");
        std::memcpy(t.vals, s + k, total);
        k += total;
        t.build();
      }
    } else if (m == 0xC0 || m == 0xC1) {  // SOF0 / SOF1
      if (sl < 6 || s[0] != 8) return fail(err, "JPEG: only 8-bit samples are supported");
      height = s[1] << 8 | s[2];
      width = s[3] << 8 | s[4];
      const int nc = s[5];
      if (width <= 0 || height <= 0) return fail(err, "JPEG: bad size");
      if (nc != 1 && nc != 3) return fail(err, "JPEG: 1 or 3 components supported");
      if (sl < (size_t)(6 + 3 * nc)) return fail(err, "JPEG: truncated SOF");
      comps.resize(nc);
      for (int c = 0; c < nc; ++c) {
        comps[c].id = s[6 + 3 * c];
        comps[c].h = s[7 + 3 * c] >> 4;
        comps[c].v = s[7 + 3 * c] & 15;
        comps[c].tq = s[8 + 3 * c] & 3;
        if (comps[c].h < 1 || comps[c].h > 4 || comps[c].v < 1 || comps[c].v > 4)
          return fail(err, "JPEG: bad sampling factors");
        hmax = std::max(hmax, comps[c].h);
        vmax = std::max(vmax, comps[c].v);
      }
      frame = true;
    } else if (m >= 0xC2 && m <= 0xCF && m != 0xC4 && m != 0xC8 && m != 0xCC) {
      return fail(err, "JPEG: progressive, lossless or arithmetic coding is not supported");
    } else if (m == 0xDD) {  // DRI
      if (sl < 2) return fail(err, "JPEG: bad DRI");
      restart = s[0] << 8 | s[1];
    } else if (m == 0xDA) {  // SOS
      if (!frame) return fail(err, "JPEG: scan before frame");
      const int ns = s[0];
      if (ns != (int)comps.size())
        return fail(err, "JPEG: scans with a subset of the components are not supported");
      for (int i = 0; i < ns; ++i) {
        const int cid = s[1 + 2 * i], tbl = s[2 + 2 * i];
        bool found = false;
        for (auto& c : comps)
          if (c.id == cid) {
            c.td = tbl >> 4;
            c.ta = tbl & 15;
            found = true;
          }
        if (!found) return fail(err, "JPEG: scan names an unknown component");
      }
      for (auto& c : comps)
        if (c.td > 3 || c.ta > 3 || !hdc[c.td].defined || !hac[c.ta].defined || !qdef[c.tq])
          return fail(err, "JPEG: missing table");
      const int mcux = (width + 8 * hmax - 1) / (8 * hmax);
      const int mcuy = (height + 8 * vmax - 1) / (8 * vmax);
      const bool single = comps.size() == 1;
      for (auto& c : comps) {
        c.bw = mcux * c.h;
        c.bh = mcuy * c.v;
        c.plane.assign((size_t)c.bw * 8 * c.bh * 8, 0);
        c.pred = 0;
      }
      Bits b{d, n, p + 2 + len};
      int coef[64];
      auto block = [&](Comp& c, int bx, int by) -> bool {
        std::memset(coef, 0, sizeof(coef));
        int t;
        if (!huff_decode(b, hdc[c.td], &t) || t > 16) return false;
        c.pred += extend(b.get(t), t);
        coef[0] = c.pred * q[c.tq][0];
        for (int k = 1; k < 64;) {
          int rs;
          if (!huff_decode(b, hac[c.ta], &rs)) return false;
          const int r = rs >> 4, sz = rs & 15;
          if (sz == 0) {
            if (r != 15) break;
            k += 16;
            continue;
          }
          k += r;
          if (k > 63) return false;
          coef[kZigzag[k]] = extend(b.get(sz), sz) * q[c.tq][k];
          ++k;
        }
        idct_block(coef, &c.plane[((size_t)by * 8 * c.bw + bx) * 8], c.bw * 8);
        return true;
      };
      // a one-component scan is not interleaved: its blocks cover the
      // component's own extent, one block per MCU
      const int bw1 = single ? (width * comps[0].h / hmax + 7) / 8 : 0;
      const int bh1 = single ? (height * comps[0].v / vmax + 7) / 8 : 0;
      const long long nmcu = single ? (long long)bw1 * bh1 : (long long)mcux * mcuy;
      for (long long mi = 0; mi < nmcu; ++mi) {
        if (restart && mi > 0 && mi % restart == 0) {
          // skip to the RSTn marker and past it, then reset the predictors
          size_t q2 = b.p;
          while (q2 + 1 < n && !(d[q2] == 0xFF && d[q2 + 1] >= 0xD0 && d[q2 + 1] <= 0xD7)) ++q2;
          if (q2 + 1 >= n) return fail(err, "JPEG: missing restart marker");
          b.p = q2 + 2;
          b.reset();
          for (auto& c : comps) c.pred = 0;
        }
        if (single) {
          if (!block(comps[0], (int)(mi % bw1), (int)(mi / bw1)))
            return fail(err, "JPEG: corrupt entropy-coded data");
        } else {
          const int mx = (int)(mi % mcux), my = (int)(mi / mcux);
          for (auto& c : comps)
            for (int v = 0; v < c.v; ++v)
              for (int h = 0; h < c.h; ++h)
                if (!block(c, mx * c.h + h, my * c.v + v))
                  return fail(err, "JPEG: corrupt entropy-coded data");
        }
      }
      // continue after the entropy-coded segment
      size_t q2 = b.p;
      while (q2 + 1 < n && !(d[q2] == 0xFF && d[q2 + 1] != 0x00 &&
                             !(d[q2 + 1] >= 0xD0 && d[q2 + 1] <= 0xD7)))
        ++q2;
      p = q2;
      done = true;  // one scan holds the whole (sequential, interleaved) image
      continue;
    }
    p += 2 + len;
  }
  if (!frame || comps.empty() || comps[0].plane.empty()) return fail(err, "JPEG: no image");
  out->width = width;
  out->height = height;
  out->rgba.assign((size_t)width * height * 4, 255);
  // full-resolution sample of component c at (x, y)
  auto upsampled_row = [&](const Comp& c, int y, std::vector<uint8_t>& row) {
    const int pw = c.bw * 8, ph = c.bh * 8;
    const int sx = hmax / c.h, sy = vmax / c.v;
    row.resize(width);
    if (sx == 1 && sy == 1) {
      std::memcpy(row.data(), &c.plane[(size_t)y * pw], width);
      return;
    }
    if ((sx == 2 && sy == 1) || (sx == 2 && sy == 2)) {
      // triangle filter: weights 3/4 near, 1/4 far in each upsampled axis
      const int cy = y / sy;
      const int cw = (width + 1) / 2;
      std::vector<int> col(cw);
      if (sy == 2) {
        const int far = (y & 1) ? std::min(cy + 1, ph - 1) : std::max(cy - 1, 0);
        for (int i = 0; i < cw; ++i)
          col[i] = 3 * c.plane[(size_t)cy * pw + i] + c.plane[(size_t)far * pw + i];
      } else {
        for (int i = 0; i < cw; ++i) col[i] = 4 * c.plane[(size_t)cy * pw + i];
      }
      for (int x = 0; x < width; ++x) {
        const int i = x >> 1;
        const int j = (x & 1) ? std::min(i + 1, cw - 1) : std::max(i - 1, 0);
        row[x] = (uint8_t)((3 * col[i] + col[j] + 8) >> 4);
      }
      return;
    }
    for (int x = 0; x < width; ++x) row[x] = c.plane[(size_t)(y / sy) * pw + x / sx];
  };
  std::vector<uint8_t> r0, r1, r2;
  for (int y = 0; y < height; ++y) {
    upsampled_row(comps[0], y, r0);
    uint8_t* o = &out->rgba[(size_t)y * width * 4];
    if (comps.size() == 1) {
      for (int x = 0; x < width; ++x) o[4 * x] = o[4 * x + 1] = o[4 * x + 2] = r0[x];
      continue;
    }
    upsampled_row(comps[1], y, r1);
    upsampled_row(comps[2], y, r2);
    for (int x = 0; x < width; ++x) {
      const float Y = r0[x], cb = (float)r1[x] - 128.f, cr = (float)r2[x] - 128.f;
      const float rgb[3] = {Y + 1.402f * cr, Y - 0.344136f * cb - 0.714136f * cr, Y + 1.772f * cb};
      for (int k = 0; k < 3; ++k)
        o[4 * x + k] = (uint8_t)std::min(255, std::max(0, (int)std::lround(rgb[k])));
    }
  }
  return true;
}

// ------------------------------------------------------------ Radiance
bool is_hdr(const uint8_t* d, size_t n) {
  return (n >= 10 && !std::memcmp(d, "#?RADIANCE", 10)) || (n >= 6 && !std::memcmp(d, "#?RGBE", 6));
}

bool decode_hdr(const uint8_t* d, size_t n, ImageF* out, std::string* err) {
  if (!is_hdr(d, n)) return fail(err, "HDR: bad signature");
  size_t p = 0;
  auto line = [&](std::string* s) {
    s->clear();
    while (p < n && d[p] != '\n') s->push_back((char)d[p++]);
    if (p < n) ++p;
    return p <= n;
  };
  std::string s;
  line(&s);
  bool format_ok = true;
  while (true) {
    if (p >= n) return fail(err, "HDR: truncated header");
    line(&s);
    if (s.empty()) break;
    if (s.rfind("FORMAT=", 0) == 0 && s != "FORMAT=32-bit_rle_rgbe") format_ok = false;
  }
  if (!format_ok) return fail(err, "HDR: only 32-bit_rle_rgbe is supported");
  line(&s);
  int w = 0, h = 0;
  if (std::sscanf(s.c_str(), "-Y %d +X %d", &h, &w) != 2 || w <= 0 || h <= 0)
    return fail(err, "HDR: only -Y h +X w orientation is supported");
  if ((long long)w * h > (1ll << 28)) return fail(err, "HDR: image too large");
  out->width = w;
  out->height = h;
  out->rgb.assign((size_t)w * h * 3, 0.f);
  std::vector<uint8_t> sc((size_t)w * 4);
  for (int y = 0; y < h; ++y) {
    if (w >= 8 && w < 32768 && p + 4 <= n && d[p] == 2 && d[p + 1] == 2 &&
        ((d[p + 2] << 8) | d[p + 3]) == w) {
      p += 4;
      for (int c = 0; c < 4; ++c) {
        int x = 0;
        while (x < w) {
          if (p >= n) return fail(err, "HDR: truncated scanline");
          int cnt = d[p++];
          if (cnt > 128) {
            cnt -= 128;
            if (p >= n || x + cnt > w) return fail(err, "HDR: bad run");
            for (int k = 0; k < cnt; ++k) sc[(size_t)(x + k) * 4 + c] = d[p];
            ++p;
          } else {
            if (cnt == 0 || x + cnt > w || p + cnt > n) return fail(err, "HDR: bad run");
            for (int k = 0; k < cnt; ++k) sc[(size_t)(x + k) * 4 + c] = d[p++];
          }
          x += cnt;
        }
      }
    } else {
      if (p + (size_t)w * 4 > n) return fail(err, "HDR: truncated scanline");
      std::memcpy(sc.data(), d + p, (size_t)w * 4);
      p += (size_t)w * 4;
    }
    for (int x = 0; x < w; ++x) {
      const uint8_t* e = &sc[(size_t)x * 4];
      float* o = &out->rgb[((size_t)y * w + x) * 3];
      if (e[3] == 0) continue;
      const float f = std::ldexp(1.0f, (int)e[3] - (128 + 8));
      o[0] = e[0] * f;
      o[1] = e[1] * f;
      o[2] = e[2] * f;
    }
  }
  return true;
}

bool decode_image8(const uint8_t* d, size_t n, Image8* out, std::string* err) {
  if (n >= 8 && d[0] == 137 && d[1] == 'P' && d[2] == 'N' && d[3] == 'G')
    return decode_png(d, n, out, err);
  if (n >= 2 && d[0] == 0xFF && d[1] == 0xD8) return decode_jpeg(d, n, out, err);
  return fail(err, "unknown image format (PNG and JPEG are supported)");
}

}  // namespace shb
