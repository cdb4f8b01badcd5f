// ORACLE — CPU restatement of the reference SH-lightmap bake path
// (DaviesX/sh-baker). TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg load it, always as the checker or as the
// timed CPU baseline, never as the product. The product (libshb_hip.so) neither
// links nor calls anything here.
//
// Every function cites the reference file:line it restates. Two random-stream
// modes share all light / material / SH / traversal code:
//   * REF  : std::mt19937(12345 + texel) with libstdc++'s distributions, in the
//            reference's draw order (baker.cpp:254, light.cpp:26-41,165,
//            material.cpp:128-130). This is the reference estimator; the GPU is
//            compared with it statistically (per-texel 3 sigma).
//   * GPU  : the counter-based stream the HIP kernels use (Philox4x32-10 per
//            path vertex seeding xoshiro128**, reservoir light pick). The GPU
//            is compared with it path-for-path (tight tolerance).
//
// Traversal: closest hit in [tnear, tfar] by the watertight FP32 test of Woop,
// Benthin, Wald (JCGT 2013) with ties resolved to the lowest triangle id, over
// a conservative binned-SAH BVH2 (or brute force); plus an FP64 brute-force
// Moller-Trumbore for checking. Embree itself is unavailable here. A second
// traversal (cw::trace) restates this build's compressed 8-wide BVH walk over
// the bytes the library exports, to recount per-ray node visits and triangle
// tests (pinned by a hand-built two-node tree with hand-counted visits).
//
// Parity pins (tests/test_oracle.py): SH basis KAT (sh_coeffs_test.cpp:25-37)
// and FP64 closed form; occlusion KATs (occlusion_test.cpp:37-59,83-111,
// 176-181) and FP64 brute force; raster KATs (rasterizer_test.cpp:26-42,65-71,
// 84-86,118-125) and an independent per-texel brute force; dilation KATs
// (dilation_test.cpp:22-32,50-57); downsample KAT (baker_test.cpp:80-87); light
// closed forms (light_test.cpp:108,196-200,226,259,274,300-305); material KATs
// (material_test.cpp:51-53,96-98,109,120); Philox and xoshiro known answers;
// the BakeSimpleQuad closed form (baker_test.cpp:10-65) in both modes; the
// bounce >= 1 estimator (Trace, vertex and texel NEE for point, spot and
// area lights, one or several light samples) against FP64 quadratures of
// plate scenes (tests/plate_closed_forms.py), each shown to fail with the
// mistake it targets planted through orc_set_fault.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <thread>
#include <vector>

#include "../include/shb.h"
#include "oracle_math.h"

namespace orc {

// ------------------------------------------------------------------ scene
struct Tex {
  uint32_t w = 0, h = 0, c = 0;
  std::vector<uint8_t> px;
};
struct Mat {
  Tex albedo;
  float emission = 0.f;
};
struct Geo {
  std::vector<V3> v, n;
  std::vector<V2> tuv, luv;
  std::vector<V4> tan;
  std::vector<uint32_t> idx;
  int mat = -1;
  float xf[12];
  uint32_t first_tri = 0;  // global id of triangle 0
};
struct Light {
  int type = 0;
  V3 pos, dir{0, 0, -1}, color{1, 1, 1};
  float intensity = 1.f, cin = 1.f, cout = 0.70710678f;
  int geom = -1, mat = -1;
};

struct BNode {
  double lo[3], hi[3];
  int left = -1, right = -1;  // internal: children
  int first = 0, count = 0;   // leaf: range in tri order
};

struct Scene {
  std::vector<Geo> geos;
  std::vector<Mat> mats;
  std::vector<Light> lights;
  // world-space triangle soup (global id = geometry-major order)
  std::vector<V3> w0, w1, w2;
  std::vector<int> tgeom, tlocal;
  std::vector<BNode> bvh;
  std::vector<int> order;  // leaf order -> global tri id
};

// src/colorspace.h:10-14
inline float SRGBToLinear(uint8_t x) {
  float v = x / 255.0f;
  if (v <= 0.04045f) return v / 12.92f;
  return std::pow((v + 0.055f) / 1.055f, 2.4f);
}

// Eigen Transform::rotation(): the unitary factor of the polar decomposition
// of the linear part, here by Newton iteration R <- (R + R^-T) / 2 in double.
static void rotation_of(const float* xf, float R[9]) {
  double m[9] = {xf[0], xf[1], xf[2], xf[4], xf[5], xf[6], xf[8], xf[9], xf[10]};
  for (int it = 0; it < 64; ++it) {
    double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6],
           h = m[7], i = m[8];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    double det = a * A + b * B + c * C;
    if (det == 0.0) break;
    // inverse transpose = cofactor / det
    double it9[9] = {A / det,
                     B / det,
                     C / det,
                     -(b * i - c * h) / det,
                     (a * i - c * g) / det,
                     -(a * h - b * g) / det,
                     (b * f - c * e) / det,
                     -(a * f - c * d) / det,
                     (a * e - b * d) / det};
    double diff = 0;
    for (int k = 0; k < 9; ++k) {
      double nv = 0.5 * (m[k] + it9[k]);
      diff = std::max(diff, std::fabs(nv - m[k]));
      m[k] = nv;
    }
    if (diff < 1e-15) break;
  }
  for (int k = 0; k < 9; ++k) R[k] = (float)m[k];
}

// Row r of a 3x3 (or 3x4) times v, reduced e0 + (e1 + e2).
static inline float rowdot(const float* row, V3 v) {
  return row[0] * v.x + (row[1] * v.y + row[2] * v.z);
}

// src/scene.cpp:162-169 TransformedVertices.
static V3 xform_point(const float* xf, V3 v) {
  return {rowdot(xf, v) + xf[3], rowdot(xf + 4, v) + xf[7],
          rowdot(xf + 8, v) + xf[11]};
}
static V3 xform_dir(const float R[9], V3 v) {
  return {rowdot(R, v), rowdot(R + 3, v), rowdot(R + 6, v)};
}

// ------------------------------------------------------------ traversal
struct RayPre {
  V3 o, d;
  int kx, ky, kz;
  float Sx, Sy, Sz;
};

static RayPre ray_pre(V3 o, V3 d) {
  RayPre r;
  r.o = o;
  r.d = d;
  float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
  int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  if (comp(d, kz) < 0.f) std::swap(kx, ky);
  r.kx = kx;
  r.ky = ky;
  r.kz = kz;
  float dz = comp(d, kz);
  r.Sx = comp(d, kx) / dz;
  r.Sy = comp(d, ky) / dz;
  r.Sz = 1.0f / dz;
  return r;
}

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013), double
// sided; u, v are the weights of vertices 1 and 2 (Embree's convention,
// occlusion.cpp:53-55).
static bool tri_isect(const RayPre& r, V3 v0, V3 v1, V3 v2, float& t,
                      float& u, float& v) {
  V3 A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
  float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
  float Ax = std::fma(-r.Sx, Akz, comp(A, r.kx));
  float Ay = std::fma(-r.Sy, Akz, comp(A, r.ky));
  float Bx = std::fma(-r.Sx, Bkz, comp(B, r.kx));
  float By = std::fma(-r.Sy, Bkz, comp(B, r.ky));
  float Cx = std::fma(-r.Sx, Ckz, comp(C, r.kx));
  float Cy = std::fma(-r.Sy, Ckz, comp(C, r.ky));
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.f || V == 0.f || W == 0.f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f))
    return false;
  float det = U + V + W;
  if (det == 0.f) return false;
  float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
  float T = std::fma(U, Az, std::fma(V, Bz, W * Cz));
  float inv = 1.0f / det;
  t = T * inv;
  u = V * inv;
  v = W * inv;
  return true;
}

static bool box_hit(const BNode& n, const V3& o, const double inv[3],
                    const bool dz[3], double tmin, double tmax) {
  const double oo[3] = {o.x, o.y, o.z};
  for (int a = 0; a < 3; ++a) {
    if (dz[a]) {
      if (oo[a] < n.lo[a] || oo[a] > n.hi[a]) return false;
      continue;
    }
    double t0 = (n.lo[a] - oo[a]) * inv[a];
    double t1 = (n.hi[a] - oo[a]) * inv[a];
    if (t0 > t1) std::swap(t0, t1);
    tmin = std::max(tmin, t0);
    tmax = std::min(tmax, t1);
    if (tmin > tmax) return false;
  }
  return true;
}

struct HitRec {
  int id = INT_MAX;
  float t = 0.f, u = 0.f, v = 0.f;
};

static thread_local uint64_t tl_node_visits = 0, tl_tri_tests = 0;

// Closest hit (or any hit) with ties to the lowest global id.
static HitRec intersect(const Scene& s, V3 o, V3 d, float tnear, float tfar,
                        bool any_hit) {
  HitRec best;
  best.t = tfar;
  if (s.w0.empty()) return best;
  RayPre r = ray_pre(o, d);
  auto test = [&](int id) {
    float t, u, v;
    ++tl_tri_tests;
    if (!tri_isect(r, s.w0[id], s.w1[id], s.w2[id], t, u, v)) return false;
    if (!(t >= tnear)) return false;
    if (t < best.t || (t == best.t && id < best.id)) {
      best.id = id;
      best.t = t;
      best.u = u;
      best.v = v;
      return true;
    }
    return false;
  };
  if (s.bvh.empty()) {
    for (int id = 0; id < (int)s.w0.size(); ++id) {
      if (test(id) && any_hit) return best;
    }
    return best;
  }
  double inv[3];
  bool dz[3] = {d.x == 0.f, d.y == 0.f, d.z == 0.f};
  inv[0] = dz[0] ? 0.0 : 1.0 / (double)d.x;
  inv[1] = dz[1] ? 0.0 : 1.0 / (double)d.y;
  inv[2] = dz[2] ? 0.0 : 1.0 / (double)d.z;
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const BNode& n = s.bvh[stack[--sp]];
    ++tl_node_visits;
    if (!box_hit(n, o, inv, dz, (double)tnear, (double)best.t)) continue;
    if (n.left < 0) {
      for (int k = 0; k < n.count; ++k) {
        if (test(s.order[n.first + k]) && any_hit) return best;
      }
    } else {
      stack[sp++] = n.right;
      stack[sp++] = n.left;
    }
  }
  return best;
}

// FP64 Moller-Trumbore brute force (checker for the FP32 watertight path).
static HitRec intersect_f64(const Scene& s, V3 o, V3 d, float tnear,
                            float tfar) {
  HitRec best;
  double bt = tfar;
  for (int id = 0; id < (int)s.w0.size(); ++id) {
    double p0[3] = {s.w0[id].x, s.w0[id].y, s.w0[id].z};
    double p1[3] = {s.w1[id].x, s.w1[id].y, s.w1[id].z};
    double p2[3] = {s.w2[id].x, s.w2[id].y, s.w2[id].z};
    double dd[3] = {d.x, d.y, d.z}, oo[3] = {o.x, o.y, o.z};
    double e1[3], e2[3], pv[3], tv[3], qv[3];
    for (int a = 0; a < 3; ++a) {
      e1[a] = p1[a] - p0[a];
      e2[a] = p2[a] - p0[a];
      tv[a] = oo[a] - p0[a];
    }
    pv[0] = dd[1] * e2[2] - dd[2] * e2[1];
    pv[1] = dd[2] * e2[0] - dd[0] * e2[2];
    pv[2] = dd[0] * e2[1] - dd[1] * e2[0];
    double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
    if (det == 0.0) continue;
    double inv = 1.0 / det;
    double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
    if (u < 0.0 || u > 1.0) continue;
    qv[0] = tv[1] * e1[2] - tv[2] * e1[1];
    qv[1] = tv[2] * e1[0] - tv[0] * e1[2];
    qv[2] = tv[0] * e1[1] - tv[1] * e1[0];
    double v = (dd[0] * qv[0] + dd[1] * qv[1] + dd[2] * qv[2]) * inv;
    if (v < 0.0 || u + v > 1.0) continue;
    double t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
    if (t < (double)tnear || t > bt) continue;
    if (t < bt || id < best.id) {
      bt = t;
      best.id = id;
      best.t = (float)t;
      best.u = (float)u;
      best.v = (float)v;
    }
  }
  return best;
}

// Binned SAH BVH2 over world triangles (stand-in for Embree's
// RTC_BUILD_QUALITY_HIGH build, scene.cpp:218). Boxes are widened by a
// relative epsilon so the FP64 slab test never culls an FP32 hit.
static void build_bvh(Scene& s) {
  const int n = (int)s.w0.size();
  s.bvh.clear();
  s.order.resize(n);
  if (n == 0) return;
  std::vector<double> blo(3 * n), bhi(3 * n), cen(3 * n);
  for (int i = 0; i < n; ++i) {
    V3 p[3] = {s.w0[i], s.w1[i], s.w2[i]};
    for (int a = 0; a < 3; ++a) {
      double lo = 1e300, hi = -1e300;
      for (int k = 0; k < 3; ++k) {
        lo = std::min(lo, (double)comp(p[k], a));
        hi = std::max(hi, (double)comp(p[k], a));
      }
      blo[3 * i + a] = lo;
      bhi[3 * i + a] = hi;
      cen[3 * i + a] = 0.5 * (lo + hi);
    }
    s.order[i] = i;
  }
  struct Item {
    int node, first, count;
  };
  s.bvh.reserve(2 * n);
  s.bvh.push_back(BNode());
  std::vector<Item> work{{0, 0, n}};
  while (!work.empty()) {
    Item it = work.back();
    work.pop_back();
    BNode& nd = s.bvh[it.node];
    double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
    for (int a = 0; a < 3; ++a) {
      nd.lo[a] = 1e300;
      nd.hi[a] = -1e300;
    }
    for (int k = it.first; k < it.first + it.count; ++k) {
      int id = s.order[k];
      for (int a = 0; a < 3; ++a) {
        nd.lo[a] = std::min(nd.lo[a], blo[3 * id + a]);
        nd.hi[a] = std::max(nd.hi[a], bhi[3 * id + a]);
        clo[a] = std::min(clo[a], cen[3 * id + a]);
        chi[a] = std::max(chi[a], cen[3 * id + a]);
      }
    }
    double ext = 0;
    for (int a = 0; a < 3; ++a) ext = std::max(ext, nd.hi[a] - nd.lo[a]);
    for (int a = 0; a < 3; ++a) {
      double pad = 1e-5 * (ext + std::fabs(nd.lo[a]) + std::fabs(nd.hi[a])) + 1e-30;
      nd.lo[a] -= pad;
      nd.hi[a] += pad;
    }
    nd.first = it.first;
    nd.count = it.count;
    if (it.count <= 4) continue;
    const int NB = 16;
    int best_axis = -1, best_split = -1;
    double best_cost = 1e300;
    for (int a = 0; a < 3; ++a) {
      double span = chi[a] - clo[a];
      if (span <= 0) continue;
      int cnt[NB] = {0};
      double bl[NB][3], bh[NB][3];
      for (int b = 0; b < NB; ++b)
        for (int q = 0; q < 3; ++q) {
          bl[b][q] = 1e300;
          bh[b][q] = -1e300;
        }
      for (int k = it.first; k < it.first + it.count; ++k) {
        int id = s.order[k];
        int b = std::min(NB - 1, (int)((cen[3 * id + a] - clo[a]) / span * NB));
        cnt[b]++;
        for (int q = 0; q < 3; ++q) {
          bl[b][q] = std::min(bl[b][q], blo[3 * id + q]);
          bh[b][q] = std::max(bh[b][q], bhi[3 * id + q]);
        }
      }
      auto area = [](const double* lo, const double* hi) {
        double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0) return 0.0;
        return dx * dy + dy * dz + dz * dx;
      };
      double la[NB], ra[NB];
      int lc[NB], rc[NB];
      double l[3] = {1e300, 1e300, 1e300}, h[3] = {-1e300, -1e300, -1e300};
      int c = 0;
      for (int b = 0; b < NB; ++b) {
        c += cnt[b];
        for (int q = 0; q < 3; ++q) {
          l[q] = std::min(l[q], bl[b][q]);
          h[q] = std::max(h[q], bh[b][q]);
        }
        la[b] = area(l, h);
        lc[b] = c;
      }
      for (int q = 0; q < 3; ++q) {
        l[q] = 1e300;
        h[q] = -1e300;
      }
      c = 0;
      for (int b = NB - 1; b >= 0; --b) {
        c += cnt[b];
        for (int q = 0; q < 3; ++q) {
          l[q] = std::min(l[q], bl[b][q]);
          h[q] = std::max(h[q], bh[b][q]);
        }
        ra[b] = area(l, h);
        rc[b] = c;
      }
      for (int b = 0; b < NB - 1; ++b) {
        if (lc[b] == 0 || rc[b + 1] == 0) continue;
        double cost = la[b] * lc[b] + ra[b + 1] * rc[b + 1];
        if (cost < best_cost) {
          best_cost = cost;
          best_axis = a;
          best_split = b;
        }
      }
    }
    int mid;
    if (best_axis < 0) {
      mid = it.first + it.count / 2;
    } else {
      double span = chi[best_axis] - clo[best_axis];
      auto pivot = std::partition(
          s.order.begin() + it.first, s.order.begin() + it.first + it.count,
          [&](int id) {
            int b = std::min(NB - 1, (int)((cen[3 * id + best_axis] - clo[best_axis]) /
                                           span * NB));
            return b <= best_split;
          });
      mid = (int)(pivot - s.order.begin());
      if (mid == it.first || mid == it.first + it.count)
        mid = it.first + it.count / 2;
    }
    int li = (int)s.bvh.size();
    s.bvh.push_back(BNode());
    int ri = (int)s.bvh.size();
    s.bvh.push_back(BNode());
    s.bvh[it.node].left = li;
    s.bvh[it.node].right = ri;
    work.push_back({ri, mid, it.first + it.count - mid});
    work.push_back({li, it.first, mid - it.first});
  }
}

// -------------------------------------------------------------- occlusion
struct Occ {
  V3 position, normal;
  V2 uv;
  int material_id = -1;
  int prim = -1, geom = -1;
  float t = 0, u = 0, v = 0;
};

// src/occlusion.cpp:9-107 FindOcclusion.
static bool FindOcclusion(const Scene& s, V3 o, V3 d, float tnear, float tfar,
                          Occ* out, uint64_t* rays) {
  if (rays) ++*rays;
  HitRec h = intersect(s, o, d, tnear, tfar, out == nullptr);
  if (h.id == INT_MAX) return false;
  if (!out) return true;
  const Geo& g = s.geos[s.tgeom[h.id]];
  int lt = s.tlocal[h.id];
  Occ occ;
  occ.t = h.t;
  occ.u = h.u;
  occ.v = h.v;
  occ.prim = h.id;
  occ.geom = s.tgeom[h.id];
  occ.position = o + d * h.t;  // occlusion.cpp:43
  occ.material_id = g.mat;     // occlusion.cpp:50
  float u = h.u, v = h.v, w = 1.0f - u - v;
  uint32_t i0 = g.idx[3 * lt], i1 = g.idx[3 * lt + 1], i2 = g.idx[3 * lt + 2];
  if (!g.n.empty()) {  // occlusion.cpp:63-69 (object-space normals)
    occ.normal = normalized(w * g.n[i0] + u * g.n[i1] + v * g.n[i2]);
  } else {  // occlusion.cpp:78-83: geometric normal, (v1-v0) x (v2-v0)
    occ.normal = normalized(cross(s.w1[h.id] - s.w0[h.id], s.w2[h.id] - s.w0[h.id]));
  }
  if (!g.tuv.empty()) {  // occlusion.cpp:86-93
    occ.uv = w * g.tuv[i0] + u * g.tuv[i1] + v * g.tuv[i2];
  }
  *out = occ;
  return true;
}

// -------------------------------------------------------------- material
// src/material.cpp:57-67
static float GetAlpha(const Mat& m, V2 uv) {
  if (m.albedo.px.empty()) return 1.0f;
  if (m.albedo.c < 4) return 1.0f;
  int tx = std::clamp((int)(uv.x * m.albedo.w), 0, (int)m.albedo.w - 1);
  int ty = std::clamp((int)(uv.y * m.albedo.h), 0, (int)m.albedo.h - 1);
  size_t idx = ((size_t)ty * m.albedo.w + tx) * m.albedo.c;
  return m.albedo.px[idx + 3] / 255.0f;
}
// src/material.cpp:69-85
static V3 GetAlbedo(const Mat& m, V2 uv) {
  V3 a{0.8f, 0.8f, 0.8f};
  if (!m.albedo.px.empty()) {
    int tx = std::clamp((int)(uv.x * m.albedo.w), 0, (int)m.albedo.w - 1);
    int ty = std::clamp((int)(uv.y * m.albedo.h), 0, (int)m.albedo.h - 1);
    size_t idx = ((size_t)ty * m.albedo.w + tx) * m.albedo.c;
    a = {SRGBToLinear(m.albedo.px[idx]), SRGBToLinear(m.albedo.px[idx + 1]),
         SRGBToLinear(m.albedo.px[idx + 2])};
  }
  return a;
}
// src/material.cpp:87-92
static V3 GetEmission(const Mat& m, V2 uv) {
  if (m.emission > 0.0f) return GetAlbedo(m, uv) * m.emission;
  return V3{};
}
// src/material.cpp:148-162 (Lambert); M_PI converted to float as Eigen does
// for a scalar quotient.
static V3 EvalMaterial(const Mat& m, V2 uv, V3 N, V3 L, V3 V) {
  if (dot(N, V) <= 0.0f || dot(N, L) <= 0.0f) return V3{};
  return GetAlbedo(m, uv) / (float)M_PI;
}
struct RefSample {
  V3 dir;
  float pdf = 0.f;
};
// src/material.cpp:113-146 given its two uniforms r1, r2.
static RefSample SampleMaterialUV(V3 n, float r1, float r2) {
  V3 t;
  if (std::fabs(n.x) > std::fabs(n.z))
    t = {-n.y, n.x, 0.0f};
  else
    t = {0.0f, -n.z, n.y};
  t = normalized(t);
  V3 b = cross(n, t);
  float phi = 2.0f * M_PI * r1;
  float theta = std::acos(std::sqrt(r2));
  float sin_theta = std::sin(theta);
  V3 l{std::cos(phi) * sin_theta, std::sin(phi) * sin_theta, std::cos(theta)};
  V3 wd = t * l.x + b * l.y + n * l.z;
  RefSample s;
  s.dir = normalized(wd);
  s.pdf = l.z / M_PI;
  return s;
}
// src/baker.cpp:38-46 given its two uniforms.
static V3 HemisphereUV(float u1, float u2) {
  float r = std::sqrt(1.0f - u1 * u1);
  float phi = 2.0f * M_PI * u2;
  return {r * std::cos(phi), r * std::sin(phi), u1};
}

// --------------------------------------------------------------- SH9
// src/sh_coeffs.cpp:50-78 (constants :27-31), direction not normalised.
static void ShBasis(V3 d, float Y[9]) {
  const float k0 = 0.282095f, k1 = 0.488603f, k21 = 1.092548f,
              k20 = 0.315392f, k22 = 0.546274f;
  float x = d.x, y = d.y, z = d.z;
  Y[0] = k0;
  Y[1] = k1 * y;
  Y[2] = k1 * z;
  Y[3] = k1 * x;
  Y[4] = k21 * x * y;
  Y[5] = k21 * y * z;
  Y[6] = k20 * (3.0f * z * z - 1.0f);
  Y[7] = k21 * x * z;
  Y[8] = k22 * (x * x - y * y);
}
static void AccumulateRadiance(V3 L, V3 d, float* sh) {
  float Y[9];
  ShBasis(d, Y);
  for (int i = 0; i < 9; ++i) {
    sh[3 * i + 0] += L.x * Y[i];
    sh[3 * i + 1] += L.y * Y[i];
    sh[3 * i + 2] += L.z * Y[i];
  }
}

// ---------------------------------------------------------------- lights
// Planted mistakes, for tests that show a check fails on them (0: none).
// Never set by anything but those tests.
enum {
  FAULT_NONE = 0,
  FAULT_CONT_NO_COSINE = 1,   // continuation weight brdf / pdf (no cos)
  FAULT_NEE_NO_COSINE = 2,    // vertex NEE radiance x brdf (no receiver cos)
  FAULT_SPOT_NO_FALLOFF = 3,  // spot cone ignored
  FAULT_AREA_PDF_TOTAL = 4    // area pdf 1 / (sum of areas) with a uniform triangle pick
};
static int g_fault = FAULT_NONE;
static inline float nee_cos(float c) { return g_fault == FAULT_NEE_NO_COSINE ? 1.0f : c; }

struct AreaSample {
  V3 point, normal, radiance;
  float pdf = 0.f;
};

// src/light.cpp:16-78 SampleAreaLight given (triangle, u1, u2).
// gpu_arith: the point as the HIP kernels compute it, v0 + u1 e1 + u2 e2
// with e1 = v1 - v0, e2 = v2 - v0 and two fused multiply-adds per axis
// (GPU-stream mode); otherwise the reference's blend.
static AreaSample SampleAreaLightAt(const Scene& s, const Light& l,
                                    size_t tri_idx, float u1, float u2,
                                    bool gpu_arith = false) {
  const Geo& geo = s.geos[l.geom];
  size_t num_triangles = geo.idx.size() / 3;
  uint32_t i0 = geo.idx[tri_idx * 3], i1 = geo.idx[tri_idx * 3 + 1],
           i2 = geo.idx[tri_idx * 3 + 2];
  if (u1 + u2 > 1.0f) {
    u1 = 1.0f - u1;
    u2 = 1.0f - u2;
  }
  float w = 1.0f - u1 - u2;
  V3 v0 = geo.v[i0], v1 = geo.v[i1], v2 = geo.v[i2];  // object space (light.cpp:45-49)
  V3 p = w * v0 + u1 * v1 + u2 * v2;
  if (gpu_arith) {
    const V3 e1 = v1 - v0, e2 = v2 - v0;
    p = V3{std::fma(u2, e2.x, std::fma(u1, e1.x, v0.x)),
           std::fma(u2, e2.y, std::fma(u1, e1.y, v0.y)),
           std::fma(u2, e2.z, std::fma(u1, e1.z, v0.z))};
  }
  V3 n{0, 1, 0};
  if (!geo.n.empty()) n = normalized(w * geo.n[i0] + u1 * geo.n[i1] + u2 * geo.n[i2]);
  V2 uv{};
  if (!geo.tuv.empty()) uv = w * geo.tuv[i0] + u1 * geo.tuv[i1] + u2 * geo.tuv[i2];
  V3 emission = GetEmission(s.mats[l.mat], uv);
  float area = norm(cross(v0 - v1, v0 - v2)) / 2.f;
  float pdf = std::max(1e-6f, 1.f / num_triangles * 1.f / area);
  if (g_fault == FAULT_AREA_PDF_TOTAL) {
    float total = 0.f;
    for (size_t t = 0; t < num_triangles; ++t) {
      V3 a = geo.v[geo.idx[3 * t]], b = geo.v[geo.idx[3 * t + 1]], c = geo.v[geo.idx[3 * t + 2]];
      total += norm(cross(a - b, a - c)) / 2.f;
    }
    pdf = std::max(1e-6f, 1.f / total);
  }
  return {p, n, emission, pdf};
}

struct ShadowRay {
  V3 o, d;
  float tnear = 0.001f, tfar = 0.f;
};

// Unshadowed incident radiance of one light (light.h:34-203) and its pending
// shadow ray. with_brdf selects the *Radiance<Brdf> form (cos * BRDF applied)
// used at path vertices; otherwise the *IncomingRadiance form used at texels.
// gpu_arith: one division per light, x * (1/d) and (1/d)^2, and the squared
// distance and N.L as x*x + (y*y + z*z) with the inner terms fused, as the
// HIP kernels compute them (GPU-stream mode); otherwise the reference's
// per-component divisions and plain sums.
static inline float sqnorm_fma(V3 a) { return std::fma(a.x, a.x, std::fma(a.y, a.y, a.z * a.z)); }
static inline float dot_fma(V3 a, V3 b) { return std::fma(a.x, b.x, std::fma(a.y, b.y, a.z * b.z)); }

static V3 LightRadiance(const Scene& s, const Light& l, const AreaSample* as,
                        V3 P, V3 N, bool with_brdf, const Mat* mat, V2 uv,
                        V3 wo, ShadowRay* ray, bool gpu_arith = false) {
  auto brdf = [&](V3 dir) { return EvalMaterial(*mat, uv, N, dir, wo); };
  switch (l.type) {
    case SHB_LIGHT_DIRECTIONAL: {  // light.h:34-64
      ray->o = P;
      ray->d = -l.dir;
      ray->tnear = 0.001f;
      ray->tfar = 1.0e10f;
      float cos_n = dot(N, -l.dir);
      if (cos_n < 0.f) return V3{};
      V3 rad = l.intensity * l.color;
      if (!with_brdf) return rad;
      return cmul(rad, nee_cos(cos_n) * brdf(-l.dir));
    }
    case SHB_LIGHT_POINT:  // light.h:72-109
    case SHB_LIGHT_SPOT: { // light.h:117-158
      V3 L = l.pos - P;
      float dist_sq = gpu_arith ? sqnorm_fma(L) : sqnorm(L);
      float dist = std::sqrt(dist_sq);
      ray->o = P;
      ray->d = l.pos - P;
      ray->tnear = 0.001f;
      ray->tfar = dist - 0.001f;
      const float inv = 1.0f / dist;
      const float inv2 = inv * inv;
      L = gpu_arith ? L * inv : L / dist;
      float cos_n = gpu_arith ? dot_fma(N, L) : dot(N, L);
      V3 rad;
      if (l.type == SHB_LIGHT_POINT) {
        if (cos_n < 0.f) return V3{};
        rad = gpu_arith ? l.intensity * l.color * inv2 : l.intensity * l.color / (dist * dist);
      } else {
        float cos_l = dot(l.dir, -L);
        float falloff = (cos_l - l.cout) / (l.cin - l.cout);
        falloff = std::clamp(falloff, 0.0f, 1.0f);
        if (g_fault == FAULT_SPOT_NO_FALLOFF) falloff = 1.0f;
        if (cos_n < 0.f) return V3{};
        rad = gpu_arith ? l.intensity * l.color * inv2 * falloff
                        : l.intensity * l.color / (dist * dist) * falloff;
      }
      if (!with_brdf) return rad;
      return cmul(rad, nee_cos(cos_n) * brdf(L));
    }
    case SHB_LIGHT_AREA: {  // light.h:166-203
      V3 L = as->point - P;
      float dist_sq = gpu_arith ? sqnorm_fma(L) : sqnorm(L);
      float dist = std::sqrt(dist_sq);
      ray->o = P;
      ray->d = as->point - P;
      ray->tnear = 0.001f;
      ray->tfar = dist - 0.001f;
      const float inv = 1.0f / dist;
      L = gpu_arith ? L * inv : L / dist;
      float cos_n = gpu_arith ? dot_fma(N, L) : dot(N, L);
      if (cos_n < 0.f) return V3{};
      V3 rad = gpu_arith ? as->radiance * (inv * inv) : as->radiance / (dist * dist);
      if (!with_brdf) return rad;
      return cmul(rad, nee_cos(cos_n) * brdf(L));
    }
    default:
      return V3{};
  }
}

static bool area_valid(const Scene& s, const Light& l) {
  return l.type == SHB_LIGHT_AREA && l.geom >= 0 && l.mat >= 0 &&
         !s.geos[l.geom].idx.empty();
}

struct Counters {
  uint64_t ext = 0, shadow = 0;
};

// ------------------------------------------------ REF stream (mt19937)
struct RefRng {
  std::mt19937& g;
  float uni() {
    std::uniform_real_distribution<float> d(0.0f, 1.0f);
    return d(g);
  }
};

// light.cpp:16-78 with the reference's draws.
static AreaSample SampleAreaLightRef(const Scene& s, const Light& l, RefRng& r) {
  if (!area_valid(s, l)) return {};
  size_t nt = s.geos[l.geom].idx.size() / 3;
  std::uniform_int_distribution<size_t> dist(0, nt - 1);
  size_t tri = dist(r.g);
  float u1 = r.uni();
  float u2 = r.uni();
  return SampleAreaLightAt(s, l, tri, u1, u2);
}

// src/light.cpp:82-189 (vertex) and 191-293 (texel, with_brdf = false).
// Returns the estimate; for the texel form accumulates SH along the
// unnormalised shadow direction (light.cpp:291).
static V3 NeeRef(const Scene& s, V3 P, V3 N, V3 wo, const Mat* mat, V2 uv,
                 unsigned num_samples, RefRng& r, bool texel, float* sh,
                 Counters& c) {
  size_t L = s.lights.size();
  std::vector<V3> rad;
  std::vector<ShadowRay> rays;
  std::vector<float> apdf, w;
  rad.reserve(L);
  rays.reserve(L);
  apdf.reserve(L);
  w.reserve(L);
  for (const Light& l : s.lights) {
    ShadowRay ray;
    float ap = 1.0f;
    V3 rd;
    if (l.type == SHB_LIGHT_AREA) {
      AreaSample as = SampleAreaLightRef(s, l, r);
      rd = LightRadiance(s, l, &as, P, N, !texel, mat, uv, wo, &ray);
      ap = as.pdf;
    } else {
      rd = LightRadiance(s, l, nullptr, P, N, !texel, mat, uv, wo, &ray);
    }
    rad.push_back(rd);
    rays.push_back(ray);
    apdf.push_back(ap);
    w.push_back(maxc(rd));
  }
  float sum = std::accumulate(w.begin(), w.end(), 0.0f);
  if (sum < 1e-6f) return V3{};
  V3 result{};
  std::discrete_distribution<int> dist(w.begin(), w.end());
  for (unsigned i = 0; i < num_samples; ++i) {
    int idx = dist(r.g);
    float pdf = w[idx] / sum;
    float joint = pdf * apdf[idx];
    // The reference traces first and tests joint_pdf after (light.cpp:173-
    // 183); the result is the same, so a dropped sample is not traced.
    if (joint < 1e-3f) continue;
    const ShadowRay& sr = rays[idx];
    if (FindOcclusion(s, sr.o, sr.d, sr.tnear, sr.tfar, nullptr, &c.shadow))
      continue;
    if (texel) {
      V3 Li = rad[idx] / (joint * num_samples);
      AccumulateRadiance(Li, sr.d, sh);
    } else {
      result += rad[idx] / joint;
    }
  }
  return result / (float)num_samples;
}

struct TraceCtx {
  const Scene* s;
  int max_depth;
  unsigned nls;
  float* vis;  // env-visibility accumulator of the texel
  Counters* c;
};

// src/baker.cpp:70-149 Trace.
static V3 TraceRef(const TraceCtx& cx, V3 origin, V3 dir, int depth, RefRng& r) {
  if (depth > cx.max_depth) return V3{};
  Occ occ;
  if (!FindOcclusion(*cx.s, origin, dir, 0.001f, INFINITY, &occ, &cx.c->ext)) {
    if (depth == 0) *cx.vis += 1.0f;
    return V3{};
  }
  const Mat& mat = cx.s->mats[occ.material_id];
  float alpha = GetAlpha(mat, occ.uv);
  V3 color{};
  if (alpha < 1.0f) {
    V3 hp = occ.position + dir * 0.001f;
    V3 tr = TraceRef(cx, hp, dir, depth + 1, r);
    color += (1.0f - alpha) * tr;
  }
  if (alpha == 0.0f) return color;
  V3 em = GetEmission(mat, occ.uv);
  if (!is_zero(em) && depth == 0) color += alpha * em;
  V3 hit_pos = occ.position + occ.normal * 0.001f;
  V3 Ld = NeeRef(*cx.s, hit_pos, occ.normal, -dir, &mat, occ.uv, cx.nls, r,
                 false, nullptr, *cx.c);
  color += alpha * Ld;
  float r1 = r.uni(), r2 = r.uni();
  RefSample smp = SampleMaterialUV(occ.normal, r1, r2);
  if (smp.pdf < 1e-3f) return color;
  V3 incoming = TraceRef(cx, hit_pos, smp.dir, depth + 1, r);
  V3 brdf = EvalMaterial(mat, occ.uv, occ.normal, smp.dir, -dir);
  float cosine = dot(occ.normal, smp.dir);
  V3 Lind = cmul(incoming, brdf) *
             ((g_fault == FAULT_CONT_NO_COSINE ? 1.0f : cosine) / smp.pdf);
  color += alpha * Lind;
  return color;
}

// ------------------------------------------ GPU stream (Philox + xoshiro)
// The stream the HIP kernels consume: every path vertex ("site") seeds a
// xoshiro128** state from Philox4x32-10(counter, key = seed). Counter:
// (texel, sample | depth << 16 | branch << 20, kind | (branch >> 12) << 8,
// job); the branch id is the vertex's node in the tree of alpha splits
// (root 1, transmitted 2b + 1, reflected 2b).
enum { SITE_TEXEL = 0, SITE_VERTEX = 1 };

static Xoshiro128ss site_rng(uint64_t seed, uint32_t texel, uint32_t sample,
                             uint32_t depth, uint32_t branch, uint32_t kind,
                             uint32_t job) {
  uint32_t ctr[4] = {texel, sample | (depth << 16) | (branch << 20),
                     kind | ((branch >> 12) << 8), job};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  philox4x32_10(ctr, key, o);
  Xoshiro128ss x;
  for (int i = 0; i < 4; ++i) x.s[i] = o[i];
  if ((o[0] | o[1] | o[2] | o[3]) == 0) x.s[0] = 1;
  return x;
}

// Unshadowed radiance, shadow ray and area pdf of light l at P in the GPU
// stream: an area light first draws its triangle (when it has several), then
// one number whose high and low 16 bits are its barycentric uniforms. With
// pick_u (the one-sample NEE) that step also gives the light's pick uniform
// (its xoshiro128+ output) and *picked is set; otherwise the caller draws
// the pick uniform itself.
static V3 GpuLightEval(const Scene& s, const Light& l, V3 P, V3 N, V3 wo, const Mat* mat,
                       V2 uv, Xoshiro128ss& x, bool texel, ShadowRay* ray, float* ap,
                       float* pick_u = nullptr, bool* picked = nullptr) {
  *ap = 1.0f;
  if (picked) *picked = false;
  if (l.type != SHB_LIGHT_AREA)
    return LightRadiance(s, l, nullptr, P, N, !texel, mat, uv, wo, ray, true);
  AreaSample as{};
  if (area_valid(s, l)) {
    uint32_t nt = (uint32_t)(s.geos[l.geom].idx.size() / 3);
    uint32_t tri = nt > 1 ? x.index(nt) : 0u;  // one triangle: no draw
    float u1, u2;
    if (pick_u) {
      x.uniform_pair16_and_pick(u1, u2, *pick_u);
      *picked = true;
    } else {
      x.uniform_pair16(u1, u2);
    }
    as = SampleAreaLightAt(s, l, tri, u1, u2, true);
  }
  V3 rd = LightRadiance(s, l, &as, P, N, !texel, mat, uv, wo, ray, true);
  *ap = as.pdf;
  return rd;
}

// NEE in the GPU stream (light.cpp:82-189 at a vertex, 191-293 at a texel).
// One sample: a reservoir pick over lights: for each light in order, its
// area draws and u (an area light with triangles: u from its pair's step,
// GpuLightEval; other lights: one uniform); W += w; pick light if u*W < w --
// light i with
// probability w_i / W, as discrete_distribution's.
// n > 1 samples: the reference's n picks with replacement from one set of
// per-light samples, as an inverse CDF: a first pass sums W, then n
// uniforms give targets u_j W, and a second pass (the same area draws,
// replayed from a copy of the state) assigns target j to the light whose
// interval of the running sum holds it. Picks are traced and summed in the
// order j; the vertex result is their sum / n, a texel adds L / (joint n)
// along each visible pick's direction.
static V3 NeeGpu(const Scene& s, V3 P, V3 N, V3 wo, const Mat* mat, V2 uv,
                 Xoshiro128ss& x, bool texel, float* sh, Counters& c, unsigned n = 1) {
  if (n > 1) {
    Xoshiro128ss x0 = x;
    float W = 0.f;
    for (const Light& l : s.lights) {
      ShadowRay ray;
      float ap;
      W += maxc(GpuLightEval(s, l, P, N, wo, mat, uv, x, texel, &ray, &ap));
    }
    std::vector<float> target(n);
    for (unsigned j = 0; j < n; ++j) target[j] = x.uniform() * W;
    if (W < 1e-6f) return V3{};
    struct Pick {
      bool found = false, ok = false;
      V3 rad;
      ShadowRay ray;
      float joint = 0.f;
    };
    std::vector<Pick> pk(n);
    float cum = 0.f;
    for (const Light& l : s.lights) {
      ShadowRay ray;
      float ap;
      V3 rd = GpuLightEval(s, l, P, N, wo, mat, uv, x0, texel, &ray, &ap);
      const float w = maxc(rd);
      const float next = cum + w;
      for (unsigned j = 0; j < n; ++j) {
        if (pk[j].found || !(w > 0.f) || !(target[j] < next)) continue;
        pk[j].found = true;
        const float joint = (w / W) * ap;
        if (joint < 1e-3f) continue;
        pk[j].ok = true;
        pk[j].rad = rd;
        pk[j].ray = ray;
        pk[j].joint = joint;
      }
      cum = next;
    }
    V3 sum{};
    for (unsigned j = 0; j < n; ++j) {
      if (!pk[j].ok) continue;
      const ShadowRay& r = pk[j].ray;
      if (FindOcclusion(s, r.o, r.d, r.tnear, r.tfar, nullptr, &c.shadow)) continue;
      if (texel)
        AccumulateRadiance(pk[j].rad / (pk[j].joint * (float)n), r.d, sh);
      else
        sum += pk[j].rad / pk[j].joint;
    }
    return texel ? V3{} : sum / (float)n;
  }
  float W = 0.f;
  int pick = -1;
  float pw = 0.f, pap = 1.f;
  V3 prad{};
  ShadowRay pray;
  for (const Light& l : s.lights) {
    ShadowRay ray;
    float ap;
    float u;
    bool picked;
    V3 rd = GpuLightEval(s, l, P, N, wo, mat, uv, x, texel, &ray, &ap, &u, &picked);
    float w = maxc(rd);
    if (!picked) u = x.uniform();
    W += w;
    if (w > 0.f && u * W < w) {
      pick = (int)(&l - s.lights.data());
      pw = w;
      pap = ap;
      prad = rd;
      pray = ray;
    }
  }
  if (W < 1e-6f || pick < 0) return V3{};
  float pdf = pw / W;
  float joint = pdf * pap;
  if (joint < 1e-3f) return V3{};
  if (FindOcclusion(s, pray.o, pray.d, pray.tnear, pray.tfar, nullptr, &c.shadow))
    return V3{};
  if (texel) {
    V3 Li = prad / (joint * 1u);
    AccumulateRadiance(Li, pray.d, sh);
    return V3{};
  }
  return (prad / joint) / 1.0f;
}

struct GpuCtx {
  const Scene* s;
  int max_depth;
  uint64_t seed;
  uint32_t texel, sample, job;
  float* vis;
  Counters* c;
  unsigned nls = 1;  // light samples per NEE
};

// Trace with the GPU stream: the control flow of TraceRef (baker.cpp:70-149)
// with counter-based numbers. The vertex at (depth, branch) draws NEE then
// the material sample from its own site. At alpha < 1 the path splits as in
// the reference: the transmitted ray, weight 1 - alpha, goes on with branch
// 2b + 1, the reflected continuation with 2b; opaque hits keep b.
static V3 TraceGpu(const GpuCtx& cx, V3 origin, V3 dir, int depth,
                   uint32_t branch) {
  if (depth > cx.max_depth) return V3{};
  Occ occ;
  if (!FindOcclusion(*cx.s, origin, dir, 0.001f, INFINITY, &occ, &cx.c->ext)) {
    if (depth == 0) *cx.vis += 1.0f;
    return V3{};
  }
  const Mat& mat = cx.s->mats[occ.material_id];
  float alpha = GetAlpha(mat, occ.uv);
  V3 color{};
  uint32_t next_branch = branch;
  if (alpha < 1.0f) {
    V3 hp = occ.position + dir * 0.001f;
    V3 tr = TraceGpu(cx, hp, dir, depth + 1, 2 * branch + 1);
    color += (1.0f - alpha) * tr;
    next_branch = 2 * branch;
  }
  if (alpha == 0.0f) return color;
  V3 em = GetEmission(mat, occ.uv);
  if (!is_zero(em) && depth == 0) color += alpha * em;
  V3 hit_pos = occ.position + occ.normal * 0.001f;
  Xoshiro128ss x = site_rng(cx.seed, cx.texel, cx.sample, (uint32_t)depth,
                            branch, SITE_VERTEX, cx.job);
  V3 Ld = NeeGpu(*cx.s, hit_pos, occ.normal, -dir, &mat, occ.uv, x, false,
                 nullptr, *cx.c, cx.nls);
  color += alpha * Ld;
  float r1 = x.uniform(), r2 = x.uniform();
  RefSample smp = SampleMaterialUV(occ.normal, r1, r2);
  if (smp.pdf < 1e-3f) return color;
  // The counter-based stream lets a zero-BRDF continuation (a back-facing
  // object-space normal) end the path without tracing it; the value is the
  // same and no later draw depends on it.
  V3 brdf = EvalMaterial(mat, occ.uv, occ.normal, smp.dir, -dir);
  if (is_zero(brdf) || depth + 1 > cx.max_depth) return color;
  V3 incoming = TraceGpu(cx, hit_pos, smp.dir, depth + 1, next_branch);
  float cosine = dot(occ.normal, smp.dir);
  V3 Lind = cmul(incoming, brdf) *
             ((g_fault == FAULT_CONT_NO_COSINE ? 1.0f : cosine) / smp.pdf);
  color += alpha * Lind;
  return color;
}

// ------------------------------------------------------------- rasterizer
// src/rasterizer.cpp:18-33
static bool Barycentric(V2 p, V2 a, V2 b, V2 c, float& u, float& v, float& w) {
  V2 v0 = b - a, v1 = c - a, v2 = p - a;
  float d00 = dot(v0, v0);
  float d01 = dot(v0, v1);
  float d11 = dot(v1, v1);
  float d20 = dot(v2, v0);
  float d21 = dot(v2, v1);
  float denom = d00 * d11 - d01 * d01;
  if (std::fabs(denom) < 1e-8f) return false;
  v = (d11 * d20 - d01 * d21) / denom;
  w = (d00 * d21 - d01 * d20) / denom;
  u = 1.0f - v - w;
  return (v >= 0.0f) && (w >= 0.0f) && (u >= 0.0f);
}

struct WorldGeo {
  std::vector<V3> v, n;
  std::vector<V4> t;
};

// src/scene.cpp:162-195
static WorldGeo world_attrs(const Geo& g) {
  WorldGeo w;
  float R[9];
  rotation_of(g.xf, R);
  for (V3 p : g.v) w.v.push_back(xform_point(g.xf, p));
  for (V3 n : g.n) w.n.push_back(normalized(xform_dir(R, n)));
  for (V4 t : g.tan) {
    V3 tv = normalized(xform_dir(R, V3{t.x, t.y, t.z}));
    w.t.push_back(V4{tv.x, tv.y, tv.z, t.w});
  }
  return w;
}

static shb_surface_point make_sp(const WorldGeo& wg, int mat, uint32_t i0,
                                 uint32_t i1, uint32_t i2, float u, float v,
                                 float w) {
  shb_surface_point sp;
  std::memset(&sp, 0, sizeof(sp));
  sp.material_id = mat;
  V3 pos = wg.v[i0] * u + wg.v[i1] * v + wg.v[i2] * w;
  V3 nrm = normalized(wg.n[i0] * u + wg.n[i1] * v + wg.n[i2] * w);
  V4 t = wg.t[i0] * u + wg.t[i1] * v + wg.t[i2] * w;
  V3 t3{t.x, t.y, t.z};
  t3 = normalized(t3 - nrm * dot(nrm, t3));
  sp.position[0] = pos.x;
  sp.position[1] = pos.y;
  sp.position[2] = pos.z;
  sp.normal[0] = nrm.x;
  sp.normal[1] = nrm.y;
  sp.normal[2] = nrm.z;
  sp.tangent[0] = t3.x;
  sp.tangent[1] = t3.y;
  sp.tangent[2] = t3.z;
  sp.tangent[3] = t.w;
  return sp;
}

}  // namespace orc

using namespace orc;

// ======================================================================
// C API for tests (ctypes). Prefix orc_.
// ======================================================================
// -------------------------------------- compressed 8-wide BVH (recount)
// Restates the traversal of this build's own compressed 8-wide BVH (no
// reference counterpart: Embree's tree is internal to the reference) to
// recount per-ray node visits and triangle tests, the figures behind the
// traversal roofline's algorithmic bytes, from the bytes shb_bvh_export
// returns. Node (80 B): origin p (3 f32), per-axis power-of-two exponents
// (3 i8), internal-child mask (u8), child base (u32), triangle base (u32),
// meta[8] (0 empty; 0x80 | rank: internal child child_base + rank; else a
// leaf of meta >> 5 triangles at tri_base + (meta & 31)), then the 8-bit
// child planes lo x, lo y, lo z, hi x, hi y, hi z (8 bytes each); child box
// plane = p + q * 2^e. Groups of children are visited nearest octant first
// (farthest first for any-hit queries, as on the device):
// bit k of a group's hit mask stands for slot k ^ sstar (sstar = the slot
// opposite the ray's direction octant); the slab test widens each plane
// distance by 16 ulps of (|b| + 256 |a|) so rounding never culls a box
// holding a hit. Triangles: 3 x float4, v0.w = the id's bits.
namespace cw {

static inline float f32_at(const uint8_t* b, int off) {
  float f;
  std::memcpy(&f, b + off, 4);
  return f;
}
static inline uint32_t u32_at(const uint8_t* b, int off) {
  uint32_t u;
  std::memcpy(&u, b + off, 4);
  return u;
}
static inline float pow2f(int e) {
  uint32_t u = (uint32_t)(e + 127) << 23;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
static inline float inv_or_big(float d) {
  if (std::fabs(d) < 1e-20f) return (d < 0.f || (d == 0.f && std::signbit(d))) ? -1e20f : 1e20f;
  return 1.0f / d;
}

struct Group {
  uint32_t base, hits, imask;
};

static void trace(const uint8_t* nodes, uint32_t nn, const float* tris, const shb_ray& ray,
                  bool any_hit, shb_hit& h, uint32_t& nv, uint32_t& tt) {
  std::memset(&h, 0, sizeof(h));
  h.prim_id = -1;
  h.geom_id = -1;
  h.material_id = -1;
  nv = tt = 0;
  const V3 o{ray.origin[0], ray.origin[1], ray.origin[2]};
  const V3 d{ray.direction[0], ray.direction[1], ray.direction[2]};
  const float sum = o.x + o.y + o.z + d.x + d.y + d.z;
  const float dd = d.x * d.x + d.y * d.y + d.z * d.z;
  if (nn == 0 || !((sum - sum) == 0.0f && dd > 0.0f && (dd - dd) == 0.0f)) return;
  const RayPre rp = ray_pre(o, d);
  const float inv[3] = {inv_or_big(d.x), inv_or_big(d.y), inv_or_big(d.z)};
  const bool par[3] = {std::fabs(d.x) < 1e-20f, std::fabs(d.y) < 1e-20f, std::fabs(d.z) < 1e-20f};
  const bool neg[3] = {d.x < 0.f, d.y < 0.f, d.z < 0.f};
  const float org[3] = {o.x, o.y, o.z};
  // nearest octant first for closest hits; an any-hit query (whose answer
  // does not depend on the order) takes the children far to near, as the
  // device traversal does
  const uint32_t sstar = ((uint32_t)neg[0] | ((uint32_t)neg[1] << 1) | ((uint32_t)neg[2] << 2)) ^
                         (any_hit ? 0u : 7u);
  float tbest = ray.tfar, bu = 0.f, bv = 0.f;
  int best = INT_MAX;
  std::vector<Group> stack;
  Group g{0u, 1u << sstar, 1u};  // the root: slot 0 of a virtual parent
  while (true) {
    if (g.hits == 0u) {
      if (stack.empty()) break;
      g = stack.back();
      stack.pop_back();
    }
    const uint32_t k = (uint32_t)__builtin_ctz(g.hits);
    g.hits &= ~(1u << k);
    const uint32_t slot = k ^ sstar;
    const uint32_t ni = g.base + (uint32_t)__builtin_popcount(g.imask & ((1u << slot) - 1u));
    if (g.hits) stack.push_back(g);
    const uint8_t* b = nodes + 80 * (size_t)ni;
    ++nv;
    float a_[3], bn[3], bf[3];
    for (int ax = 0; ax < 3; ++ax) {
      const float sc = pow2f((int)(int8_t)b[12 + ax]);
      const float a = sc * inv[ax];
      const float bb = (f32_at(b, 4 * ax) - org[ax]) * inv[ax];
      const float e = par[ax] ? 0.f : (std::fabs(bb) + 256.0f * std::fabs(a)) * 0x1p-20f;
      a_[ax] = a;
      bn[ax] = bb - e;
      bf[ax] = bb + e;
    }
    const uint8_t imask = b[15];
    const uint8_t* meta = b + 24;
    const uint8_t* lo[3] = {b + 32, b + 40, b + 48};
    const uint8_t* hi[3] = {b + 56, b + 64, b + 72};
    uint32_t hits = 0, tmask = 0;
    const float tcap = tbest;
    for (int c = 0; c < 8; ++c) {
      float tn[3], tf[3];
      for (int ax = 0; ax < 3; ++ax) {
        const uint8_t qn = neg[ax] ? hi[ax][c] : lo[ax][c];
        const uint8_t qf = neg[ax] ? lo[ax][c] : hi[ax][c];
        tn[ax] = std::fma((float)qn, a_[ax], bn[ax]);
        tf[ax] = std::fma((float)qf, a_[ax], bf[ax]);
      }
      const float tmin = std::fmax(std::fmax(tn[0], tn[1]), std::fmax(tn[2], ray.tnear));
      const float tmax = std::fmin(std::fmin(std::fmin(tf[0], tf[1]), tf[2]), tcap);
      const uint32_t m = meta[c];
      if (m != 0u && tmin <= tmax) {
        if (m & 0x80u)
          hits |= 1u << ((uint32_t)c ^ sstar);
        else
          tmask |= ((1u << (m >> 5)) - 1u) << (m & 31u);
      }
    }
    g = Group{u32_at(b, 16), hits, imask};
    const uint32_t tri_base = u32_at(b, 20);
    bool found = false;
    while (tmask) {
      const uint32_t j = (uint32_t)__builtin_ctz(tmask);
      tmask &= tmask - 1u;
      const float* tp = tris + 12 * (size_t)(tri_base + j);
      ++tt;
      float t, u, v;
      if (tri_isect(rp, V3{tp[0], tp[1], tp[2]}, V3{tp[4], tp[5], tp[6]},
                    V3{tp[8], tp[9], tp[10]}, t, u, v)) {
        uint32_t idb;
        std::memcpy(&idb, tp + 3, 4);
        const int id = (int)idb;
        if (t >= ray.tnear && (t < tbest || (t == tbest && id < best))) {
          tbest = t;
          best = id;
          bu = u;
          bv = v;
          found = true;
        }
      }
    }
    if (any_hit && found) break;
  }
  if (best != INT_MAX) {
    h.prim_id = best;
    h.t = tbest;
    h.u = bu;
    h.v = bv;
  }
}

}  // namespace cw

extern "C" {

void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  philox4x32_10(ctr, key, out);
}

void orc_xoshiro128ss(const uint32_t* state, int n, uint32_t* out) {
  Xoshiro128ss x;
  for (int i = 0; i < 4; ++i) x.s[i] = state[i];
  for (int i = 0; i < n; ++i) out[i] = x.next();
}

void orc_site_uniforms(uint64_t seed, uint32_t texel, uint32_t sample,
                       uint32_t depth, uint32_t branch, uint32_t kind,
                       uint32_t job, int n, float* out) {
  Xoshiro128ss x = site_rng(seed, texel, sample, depth, branch, kind, job);
  for (int i = 0; i < n; ++i) out[i] = x.uniform();
}

// sh_coeffs.cpp:50-78
void orc_sh_accumulate(const float* radiance, const float* dir, float* sh27) {
  AccumulateRadiance(V3{radiance[0], radiance[1], radiance[2]},
                     V3{dir[0], dir[1], dir[2]}, sh27);
}

float orc_srgb_to_linear(uint8_t x) { return SRGBToLinear(x); }

// Plants mistake `fault` (FAULT_*) in every later call; 0 removes it.
void orc_set_fault(int fault) { g_fault = fault; }

// material.cpp:113-146 with given uniforms: out = dir[3], pdf
void orc_sample_material(const float* n, float r1, float r2, float* out4) {
  RefSample s = SampleMaterialUV(V3{n[0], n[1], n[2]}, r1, r2);
  out4[0] = s.dir.x;
  out4[1] = s.dir.y;
  out4[2] = s.dir.z;
  out4[3] = s.pdf;
}

void* orc_scene_create(const shb_scene_desc* d, int use_bvh) {
  Scene* s = new Scene();
  for (uint32_t m = 0; m < d->num_materials; ++m) {
    const shb_material& sm = d->materials[m];
    Mat mat;
    mat.emission = sm.emission_intensity;
    if (sm.albedo_channels > 0) {
      mat.albedo.w = sm.albedo_width;
      mat.albedo.h = sm.albedo_height;
      mat.albedo.c = sm.albedo_channels;
      size_t nb = (size_t)sm.albedo_width * sm.albedo_height * sm.albedo_channels;
      mat.albedo.px.assign(d->texel_pool + sm.albedo_offset,
                           d->texel_pool + sm.albedo_offset + nb);
    }
    s->mats.push_back(std::move(mat));
  }
  uint32_t tri_base = 0;
  for (uint32_t gi = 0; gi < d->num_geometries; ++gi) {
    const shb_geometry& sg = d->geometries[gi];
    Geo g;
    g.mat = sg.material_id;
    std::memcpy(g.xf, sg.transform, sizeof(g.xf));
    for (uint32_t k = 0; k < sg.vertex_count; ++k) {
      uint32_t vi = sg.vertex_offset + k;
      g.v.push_back(V3{d->positions[3 * vi], d->positions[3 * vi + 1],
                       d->positions[3 * vi + 2]});
      if ((sg.flags & SHB_GEOM_HAS_NORMALS) && d->normals)
        g.n.push_back(V3{d->normals[3 * vi], d->normals[3 * vi + 1],
                         d->normals[3 * vi + 2]});
      if ((sg.flags & SHB_GEOM_HAS_TEXTURE_UVS) && d->texture_uvs)
        g.tuv.push_back(V2{d->texture_uvs[2 * vi], d->texture_uvs[2 * vi + 1]});
      if ((sg.flags & SHB_GEOM_HAS_LIGHTMAP_UVS) && d->lightmap_uvs)
        g.luv.push_back(V2{d->lightmap_uvs[2 * vi], d->lightmap_uvs[2 * vi + 1]});
      if ((sg.flags & SHB_GEOM_HAS_TANGENTS) && d->tangents)
        g.tan.push_back(V4{d->tangents[4 * vi], d->tangents[4 * vi + 1],
                           d->tangents[4 * vi + 2], d->tangents[4 * vi + 3]});
    }
    g.idx.assign(d->indices + sg.index_offset,
                 d->indices + sg.index_offset + (sg.index_count / 3) * 3);
    g.first_tri = tri_base;
    // scene.cpp:209-262: world-space vertex copy per geometry.
    std::vector<V3> wv;
    for (V3 p : g.v) wv.push_back(xform_point(g.xf, p));
    uint32_t nt = (uint32_t)(g.idx.size() / 3);
    for (uint32_t t = 0; t < nt; ++t) {
      s->w0.push_back(wv[g.idx[3 * t]]);
      s->w1.push_back(wv[g.idx[3 * t + 1]]);
      s->w2.push_back(wv[g.idx[3 * t + 2]]);
      s->tgeom.push_back((int)gi);
      s->tlocal.push_back((int)t);
    }
    tri_base += nt;
    s->geos.push_back(std::move(g));
  }
  for (uint32_t li = 0; li < d->num_lights; ++li) {
    const shb_light& sl = d->lights[li];
    Light l;
    l.type = sl.type;
    l.pos = V3{sl.position[0], sl.position[1], sl.position[2]};
    l.dir = V3{sl.direction[0], sl.direction[1], sl.direction[2]};
    l.color = V3{sl.color[0], sl.color[1], sl.color[2]};
    l.intensity = sl.intensity;
    l.cin = sl.cos_inner_cone;
    l.cout = sl.cos_outer_cone;
    l.geom = sl.geometry_index;
    l.mat = sl.material_index;
    s->lights.push_back(l);
  }
  if (use_bvh) build_bvh(*s);
  return s;
}

void orc_scene_destroy(void* s) { delete (Scene*)s; }

// FindOcclusion over a ray list. mode 0: FP32 watertight (BVH if built),
// 1: FP64 brute force, 2: FP32 any-hit. Optional per-ray counters.
// Traversal of an exported compressed 8-wide BVH (see cw::trace) over a ray
// list: closest hit (any_hit 0) or any hit; per-ray node visits and
// triangle tests.
void orc_cwbvh_trace(const uint8_t* nodes, uint32_t num_nodes, const float* tris,
                     const shb_ray* rays, uint32_t n, int any_hit, shb_hit* hits,
                     uint32_t* node_visits, uint32_t* tri_tests) {
  for (uint32_t i = 0; i < n; ++i)
    cw::trace(nodes, num_nodes, tris, rays[i], any_hit != 0, hits[i], node_visits[i],
              tri_tests[i]);
}

void orc_trace_rays(void* sp, const shb_ray* rays, uint32_t n, int mode,
                    shb_hit* hits, uint32_t* node_visits, uint32_t* tri_tests) {
  const Scene& s = *(Scene*)sp;
  for (uint32_t i = 0; i < n; ++i) {
    const shb_ray& r = rays[i];
    V3 o{r.origin[0], r.origin[1], r.origin[2]};
    V3 d{r.direction[0], r.direction[1], r.direction[2]};
    shb_hit h;
    std::memset(&h, 0, sizeof(h));
    h.prim_id = -1;
    h.geom_id = -1;
    h.material_id = -1;
    tl_node_visits = 0;
    tl_tri_tests = 0;
    if (mode == 1) {
      HitRec hr = intersect_f64(s, o, d, r.tnear, r.tfar);
      if (hr.id != INT_MAX) {
        h.prim_id = hr.id;
        h.t = hr.t;
        h.u = hr.u;
        h.v = hr.v;
        h.geom_id = s.tgeom[hr.id];
        h.material_id = s.geos[h.geom_id].mat;
      }
    } else if (mode == 2) {
      HitRec hr = intersect(s, o, d, r.tnear, r.tfar, true);
      if (hr.id != INT_MAX) h.prim_id = hr.id;
    } else {
      Occ occ;
      if (FindOcclusion(s, o, d, r.tnear, r.tfar, &occ, nullptr)) {
        h.t = occ.t;
        h.u = occ.u;
        h.v = occ.v;
        h.prim_id = occ.prim;
        h.geom_id = occ.geom;
        h.material_id = occ.material_id;
        h.position[0] = occ.position.x;
        h.position[1] = occ.position.y;
        h.position[2] = occ.position.z;
        h.normal[0] = occ.normal.x;
        h.normal[1] = occ.normal.y;
        h.normal[2] = occ.normal.z;
        h.uv[0] = occ.uv.x;
        h.uv[1] = occ.uv.y;
      }
    }
    if (node_visits) node_visits[i] = (uint32_t)tl_node_visits;
    if (tri_tests) tri_tests[i] = (uint32_t)tl_tri_tests;
    hits[i] = h;
  }
}

// material.cpp:57-92: out = albedo[3], alpha, emission[3]
void orc_material_eval(void* sp, int mat, const float* uv, float* out7) {
  const Scene& s = *(Scene*)sp;
  V2 u{uv[0], uv[1]};
  V3 a = GetAlbedo(s.mats[mat], u);
  V3 e = GetEmission(s.mats[mat], u);
  out7[0] = a.x;
  out7[1] = a.y;
  out7[2] = a.z;
  out7[3] = GetAlpha(s.mats[mat], u);
  out7[4] = e.x;
  out7[5] = e.y;
  out7[6] = e.z;
}

// light.h:34-203 single light, given brdf = constant (1,1,1) when
// with_brdf < 0 (the tests' simple_brdf), material brdf when 1, none when 0.
// area_in = (tri, u1, u2) for area lights. out = radiance[3], ray o,d,tnear,tfar
void orc_light_radiance(void* sp, int li, const float* P, const float* N,
                        const float* wo, int mat, const float* uv, int with_brdf,
                        const float* area_in, float* out11) {
  const Scene& s = *(Scene*)sp;
  const Light& l = s.lights[li];
  V3 p{P[0], P[1], P[2]}, n{N[0], N[1], N[2]}, w{wo[0], wo[1], wo[2]};
  V2 u{uv[0], uv[1]};
  ShadowRay ray;
  AreaSample as{};
  if (l.type == SHB_LIGHT_AREA)
    as = SampleAreaLightAt(s, l, (size_t)area_in[0], area_in[1], area_in[2]);
  V3 r;
  if (with_brdf < 0) {
    // constant unit BRDF: radiance * cos_n
    r = LightRadiance(s, l, &as, p, n, false, nullptr, u, w, &ray);
    float cos_n;
    if (l.type == SHB_LIGHT_DIRECTIONAL)
      cos_n = dot(n, -l.dir);
    else
      cos_n = dot(n, normalized(ray.d));
    if (cos_n < 0.f) r = V3{};
    else r = cmul(r, cos_n * V3{1, 1, 1});
  } else {
    r = LightRadiance(s, l, &as, p, n, with_brdf != 0,
                      mat >= 0 ? &s.mats[mat] : nullptr, u, w, &ray);
  }
  out11[0] = r.x;
  out11[1] = r.y;
  out11[2] = r.z;
  out11[3] = ray.o.x;
  out11[4] = ray.o.y;
  out11[5] = ray.o.z;
  out11[6] = ray.d.x;
  out11[7] = ray.d.y;
  out11[8] = ray.d.z;
  out11[9] = ray.tnear;
  out11[10] = ray.tfar;
}

// light.cpp:16-78: out = point[3], normal[3], radiance[3], pdf
void orc_sample_area_light(void* sp, int li, uint32_t tri, float u1, float u2,
                           float* out10) {
  const Scene& s = *(Scene*)sp;
  AreaSample a = SampleAreaLightAt(s, s.lights[li], tri, u1, u2);
  float v[10] = {a.point.x,    a.point.y,    a.point.z,    a.normal.x,
                 a.normal.y,   a.normal.z,   a.radiance.x, a.radiance.y,
                 a.radiance.z, a.pdf};
  std::memcpy(out10, v, sizeof(v));
}

// light.cpp:82-189 EvaluateLightSamples with std::mt19937(seed).
void orc_evaluate_light_samples(void* sp, const float* P, const float* N,
                                const float* wo, int mat, const float* uv,
                                unsigned num_samples, uint32_t seed,
                                float* out3) {
  const Scene& s = *(Scene*)sp;
  std::mt19937 g(seed);
  RefRng r{g};
  Counters c;
  V3 res = NeeRef(s, V3{P[0], P[1], P[2]}, V3{N[0], N[1], N[2]},
                  V3{wo[0], wo[1], wo[2]}, &s.mats[mat], V2{uv[0], uv[1]},
                  num_samples, r, false, nullptr, c);
  out3[0] = res.x;
  out3[1] = res.y;
  out3[2] = res.z;
}

// src/rasterizer.cpp:68-211 in serial order (geometries, then triangles).
void orc_rasterize(void* sp, int width, int height, int ss,
                   shb_surface_point* out) {
  const Scene& s = *(Scene*)sp;
  int W = width * ss, H = height * ss;
  for (int i = 0; i < W * H; ++i) {
    std::memset(&out[i], 0, sizeof(shb_surface_point));
    out[i].material_id = -1;
  }
  for (const Geo& g : s.geos) {
    if (g.luv.empty() || g.n.empty() || g.tan.empty()) continue;
    WorldGeo wg = world_attrs(g);
    size_t nt = g.idx.size() / 3;
    for (size_t i = 0; i < nt; ++i) {
      uint32_t i0 = g.idx[3 * i], i1 = g.idx[3 * i + 1], i2 = g.idx[3 * i + 2];
      V2 a = g.luv[i0], b = g.luv[i1], c = g.luv[i2];
      float min_u = std::min({a.x, b.x, c.x}), max_u = std::max({a.x, b.x, c.x});
      float min_v = std::min({a.y, b.y, c.y}), max_v = std::max({a.y, b.y, c.y});
      int min_x = std::max(0, (int)(min_u * W));
      int max_x = std::min(W - 1, (int)(std::ceil(max_u * W)));
      int min_y = std::max(0, (int)(min_v * H));
      int max_y = std::min(H - 1, (int)(std::ceil(max_v * H)));
      bool any = false;
      for (int y = min_y; y <= max_y; ++y) {
        for (int x = min_x; x <= max_x; ++x) {
          V2 p{(x + 0.5f) / W, (y + 0.5f) / H};
          float u, v, w;
          if (!Barycentric(p, a, b, c, u, v, w)) continue;
          any = true;
          out[(size_t)y * W + x] = make_sp(wg, g.mat, i0, i1, i2, u, v, w);
        }
      }
      if (!any) {  // rasterizer.cpp:160-205
        float cu = (a.x + b.x + c.x) / 3.0f;
        float cv = (a.y + b.y + c.y) / 3.0f;
        int cx = (int)(cu * W), cy = (int)(cv * H);
        if (cx >= 0 && cx < W && cy >= 0 && cy < H) {
          size_t pi = (size_t)cy * W + cx;
          if (out[pi].material_id < 0) {
            float t3 = 1.0f / 3.0f;
            out[pi] = make_sp(wg, g.mat, i0, i1, i2, t3, t3, t3);
          }
        }
      }
    }
  }
}

// rasterizer.cpp:213-220
void orc_validity_mask(const shb_surface_point* pts, size_t n, uint8_t* mask) {
  for (size_t i = 0; i < n; ++i) mask[i] = pts[i].material_id >= 0;
}

// rasterizer.cpp:222-241 (width, height: output size)
void orc_downsample_mask(const uint8_t* in, int width, int height, int scale,
                         uint8_t* out) {
  int stride = width * scale;
  for (int y = 0; y < height; ++y)
    for (int x = 0; x < width; ++x) {
      bool any = false;
      for (int sy = y * scale; sy < (y + 1) * scale && !any; ++sy)
        for (int sx = x * scale; sx < (x + 1) * scale; ++sx)
          if (in[(size_t)sy * stride + sx]) {
            any = true;
            break;
          }
      out[(size_t)y * width + x] = any;
    }
}

// baker.cpp:151-182 DownsampleTexture for `ch` floats per texel; in_w, in_h:
// input size.
void orc_downsample(const float* in, int in_w, int in_h, int scale, int ch,
                    float* out) {
  int ow = in_w / scale, oh = in_h / scale;
  std::vector<float> avg(ch);
  for (int y = 0; y < oh; ++y)
    for (int x = 0; x < ow; ++x) {
      std::fill(avg.begin(), avg.end(), 0.f);
      int count = 0;
      for (int dy = 0; dy < scale; ++dy)
        for (int dx = 0; dx < scale; ++dx) {
          int sx = x * scale + dx, sy = y * scale + dy;
          if (sx < in_w && sy < in_h) {
            const float* p = in + ((size_t)sy * in_w + sx) * ch;
            for (int c = 0; c < ch; ++c) avg[c] += p[c];
            count++;
          }
        }
      float* o = out + ((size_t)y * ow + x) * ch;
      if (count > 0) {
        float r = 1.f / float(count);
        for (int c = 0; c < ch; ++c) o[c] = avg[c] * r;
      } else {
        for (int c = 0; c < ch; ++c) o[c] = 0.f;
      }
    }
}

// dilation.cpp:7-73 for `ch` floats per texel (SHCoeffs: ch = 27; float: 1).
void orc_dilate(int width, int height, int ch, float* pixels, uint8_t* mask,
                int passes) {
  size_t n = (size_t)width * height;
  if (n == 0) return;
  std::vector<float> next(pixels, pixels + n * ch);
  std::vector<uint8_t> next_mask(mask, mask + n);
  std::vector<float> acc(ch);
  for (int p = 0; p < passes; ++p) {
    bool changed = false;
    for (int y = 0; y < height; ++y)
      for (int x = 0; x < width; ++x) {
        size_t idx = (size_t)y * width + x;
        if (mask[idx]) continue;
        std::fill(acc.begin(), acc.end(), 0.f);
        int count = 0;
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            if (dx == 0 && dy == 0) continue;
            int nx = x + dx, ny = y + dy;
            if (nx >= 0 && nx < width && ny >= 0 && ny < height) {
              size_t ni = (size_t)ny * width + nx;
              if (mask[ni]) {
                for (int c = 0; c < ch; ++c) acc[c] += pixels[ni * ch + c];
                count++;
              }
            }
          }
        if (count > 0) {
          float r = 1.0f / count;
          for (int c = 0; c < ch; ++c) next[idx * ch + c] = acc[c] * r;
          next_mask[idx] = 1;
          changed = true;
        }
      }
    std::memcpy(pixels, next.data(), n * ch * sizeof(float));
    std::memcpy(mask, next_mask.data(), n);
    if (!changed) break;
  }
}

struct OrcBakeStats {
  uint64_t valid_texels, paths, rays_extension, rays_shadow;
  double seconds;
  int threads;
};

// src/baker.cpp:186-304 BakeSHLightMap over (W*ss) x (H*ss) surface points.
// rng_mode 0: reference stream; 1: GPU stream. Only rows y % row_stride == 0
// are baked (row_stride 1 = all), for bounded CPU-baseline samples.
// var_out (optional, 27 floats per texel) receives the per-texel sample
// variance of the SH estimate's mean (for the 3-sigma test).
int orc_bake(void* sp, const shb_surface_point* pts, int width, int height,
             int samples, int bounces, unsigned num_light_samples, int rng_mode,
             uint64_t seed, uint32_t job, int threads, int row_stride,
             float* sh_out, float* vis_out, float* var_out,
             OrcBakeStats* stats) {
  const Scene& s = *(Scene*)sp;
  auto t0 = std::chrono::steady_clock::now();
  const size_t total = (size_t)width * height;
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  if (row_stride < 1) row_stride = 1;
  std::atomic<size_t> next{0};
  std::atomic<uint64_t> ext{0}, shadow{0}, valid{0};
  const float inv_pdf_uniform = 2.0f * M_PI;  // baker.cpp:221
  auto worker = [&]() {
    Counters c;
    uint64_t nvalid = 0;
    std::vector<float> sh(27), nee1(27), ind1(27);
    // per-sample sums for the variance, in double apart from the estimate
    // itself (a float sum of squares loses the variance of nearly constant
    // samples to cancellation)
    std::vector<double> s1(27), s2(27);
    for (;;) {
      size_t begin = next.fetch_add(256);
      if (begin >= total) break;
      size_t end = std::min(total, begin + 256);
      for (size_t idx = begin; idx < end; ++idx) {
        float* out = sh_out + idx * 27;
        std::fill(out, out + 27, 0.f);
        if (vis_out) vis_out[idx] = 0.f;
        if (var_out) std::fill(var_out + idx * 27, var_out + idx * 27 + 27, 0.f);
        const shb_surface_point& p = pts[idx];
        if (p.material_id < 0) continue;
        if ((int)(idx / width) % row_stride != 0) continue;
        ++nvalid;
        std::fill(sh.begin(), sh.end(), 0.f);
        std::fill(s1.begin(), s1.end(), 0.0);
        std::fill(s2.begin(), s2.end(), 0.0);
        float vis = 0.f;
        V3 P{p.position[0], p.position[1], p.position[2]};
        V3 N{p.normal[0], p.normal[1], p.normal[2]};
        V3 T{p.tangent[0], p.tangent[1], p.tangent[2]};
        float Tw = p.tangent[3];
        V3 origin = P + N * 0.001f;  // baker.cpp:257-259
        std::mt19937 g(12345 + idx);  // baker.cpp:254
        RefRng rr{g};
        for (int smp = 0; smp < samples; ++smp) {
          std::fill(nee1.begin(), nee1.end(), 0.f);
          std::fill(ind1.begin(), ind1.end(), 0.f);
          V3 dl;
          Xoshiro128ss x{};
          if (rng_mode == 0) {
            float u1 = rr.uni(), u2 = rr.uni();
            dl = HemisphereUV(u1, u2);
          } else {
            x = site_rng(seed, (uint32_t)idx, (uint32_t)smp, 0, 1, SITE_TEXEL, job);
            float u1 = x.uniform(), u2 = x.uniform();
            dl = HemisphereUV(u1, u2);
          }
          V3 bit = cross(N, T) * Tw;  // baker.cpp:267-271
          V3 dw = T * dl.x + bit * dl.y + N * dl.z;
          V3 Li;
          if (rng_mode == 0) {
            NeeRef(s, P, N, V3{}, nullptr, V2{}, num_light_samples, rr, true,
                   nee1.data(), c);
            TraceCtx cx{&s, bounces, num_light_samples, &vis, &c};
            Li = TraceRef(cx, origin, dw, 0, rr) * inv_pdf_uniform;
          } else {
            NeeGpu(s, P, N, V3{}, nullptr, V2{}, x, true, nee1.data(), c, num_light_samples);
            GpuCtx cx{&s, bounces, seed, (uint32_t)idx, (uint32_t)smp, job, &vis, &c,
                      num_light_samples};
            Li = TraceGpu(cx, origin, dw, 0, 1) * inv_pdf_uniform;
          }
          AccumulateRadiance(Li, dw, ind1.data());
          // sh += NEE term, then sh += indirect term (baker.cpp:274-288);
          // with one light sample each term is a single L*Y product.
          for (int k = 0; k < 27; ++k) {
            sh[k] += nee1[k];
            sh[k] += ind1[k];
            const double one = (double)nee1[k] + (double)ind1[k];
            s1[k] += one;
            s2[k] += one * one;
          }
        }
        float inv = 1.0f / samples;  // baker.cpp:292-294
        for (int k = 0; k < 27; ++k) out[k] = sh[k] * inv;
        if (vis_out) vis_out[idx] = vis * inv;
        if (var_out && samples > 1) {
          for (int k = 0; k < 27; ++k) {
            const double m = s1[k] / samples;
            const double v = (s2[k] / samples - m * m) * samples / (samples - 1);
            var_out[idx * 27 + k] = (float)(std::max(0.0, v) / samples);
          }
        }
      }
    }
    ext += c.ext;
    shadow += c.shadow;
    valid += nvalid;
  };
  std::vector<std::thread> pool;
  for (int i = 0; i < threads; ++i) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
  auto t1 = std::chrono::steady_clock::now();
  if (stats) {
    stats->valid_texels = valid;
    stats->paths = valid * (uint64_t)samples;
    stats->rays_extension = ext;
    stats->rays_shadow = shadow;
    stats->seconds = std::chrono::duration<double>(t1 - t0).count();
    stats->threads = threads;
  }
  return 0;
}

// ------------------------------------------------------- environment SH
// src/scene.cpp:31-71 ProjectLatLongMap: the reference's serial float sum
// into sh_ref, and the FP64 sum of the same float terms into sh64.
void orc_env_latlong(const float* px, int w, int h, int ch, float* sh_ref, double* sh64) {
  const float kPiF = 3.14159265359f, kTwoPiF = 6.28318530718f;  // scene.cpp:12-13
  const float d_phi = kTwoPiF / w, d_theta = kPiF / h;
  std::fill(sh_ref, sh_ref + 27, 0.0f);
  std::fill(sh64, sh64 + 27, 0.0);
  for (int y = 0; y < h; ++y) {
    const float theta = (y + 0.5f) * d_theta;
    const float sin_theta = std::sin(theta);
    const float d_omega = d_phi * d_theta * sin_theta;
    if (sin_theta < 1e-4f) continue;
    for (int x = 0; x < w; ++x) {
      const float phi = (x + 0.5f) * d_phi;
      const V3 dir{std::sin(theta) * std::sin(phi), std::cos(theta),
                   std::sin(theta) * std::cos(phi)};
      const size_t idx = ((size_t)y * w + x) * ch;
      const V3 L = V3{px[idx], px[idx + 1], px[idx + 2]} * d_omega;
      AccumulateRadiance(L, dir, sh_ref);
      float Y[9];
      ShBasis(dir, Y);
      for (int i = 0; i < 9; ++i) {
        sh64[3 * i + 0] += (double)(L.x * Y[i]);
        sh64[3 * i + 1] += (double)(L.y * Y[i]);
        sh64[3 * i + 2] += (double)(L.z * Y[i]);
      }
    }
  }
}

// src/scene.cpp:75-94 EvaluatePerez, coefficients :15-20.
static float EvaluatePerez(V3 dir, V3 sun) {
  const float A = -1.0f, B = -0.32f, C = 10.0f, D = -3.0f, E = 0.45f;
  const float cos_zenith = std::max(0.001f, dir.y);
  const float cos_sun = dot(dir, sun);
  const float gamma = std::acos(std::clamp(cos_sun, -1.0f, 1.0f));
  const float f_zenith = (1.0f + A * std::exp(B / cos_zenith));
  const float f_sun = (1.0f + C * std::exp(D * gamma) + E * cos_sun * cos_sun);
  return f_zenith * f_sun;
}

// src/scene.cpp:96-150 ProjectPreethamToSH.
void orc_env_preetham(const float* sun_direction, float* sh) {
  const float kPiF = 3.14159265359f, kTwoPiF = 6.28318530718f;
  const int steps_theta = 64, steps_phi = 128;
  const float d_theta = kPiF / steps_theta, d_phi = kTwoPiF / steps_phi;
  V3 s{sun_direction[0], sun_direction[1], sun_direction[2]};
  const float n2 = dot(s, s);
  const V3 sun = n2 > 0.0f ? s / std::sqrt(n2) : s;  // Eigen normalized()
  float zenith = EvaluatePerez(V3{0, 1, 0}, sun);
  if (zenith <= 1e-6f) zenith = 1.0f;
  std::fill(sh, sh + 27, 0.0f);
  for (int t = 0; t < steps_theta; ++t) {
    const float theta = (t + 0.5f) * d_theta;
    const float sin_theta = std::sin(theta);
    for (int p = 0; p < steps_phi; ++p) {
      const float phi = (p + 0.5f) * d_phi;
      const float y = std::cos(theta), x = sin_theta * std::sin(phi),
                  z = sin_theta * std::cos(phi);
      const V3 dir{x, y, z};
      if (y < 0.0f) continue;
      const float lum = std::min(EvaluatePerez(dir, sun), zenith * 30.0f);
      const V3 rad = V3{0.2f, 0.5f, 0.9f} * (lum / zenith);
      AccumulateRadiance(rad * (sin_theta * d_theta * d_phi), dir, sh);
    }
  }
}

}  // extern "C"
