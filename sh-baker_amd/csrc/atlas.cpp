// Lightmap atlas (host), replacing CreateAtlasGeometries and
// atlas_internal::CalculateGeometryScales (src/atlas.cpp:22-250), which hand
// the scaled world-space meshes to xatlas. xatlas is not available here, so
// the charting and packing are this file's own:
//
//  * per-geometry scale: the reference's heuristic, restated exactly
//    (density_multiplier * sqrt(albedo w * h) * sqrt(UV tile count), clamped
//    to 5x the median);
//  * charts: faces are grown from the largest unassigned face over edges
//    shared by welded positions, while the face normal stays within 25
//    degrees of both the seed normal and the chart's running area-weighted
//    normal; the projection axis is the final mean normal, or the seed
//    normal when the mean drifted (every face within 25 degrees of it);
//  * parameterisation: orthographic projection onto the plane orthogonal
//    to the chart normal, rotated to the minimum-area bounding rectangle of
//    its convex hull (rotating calipers), so world-space texel density is
//    kept within cos 25 degrees;
//  * packing: rectangles shelf-packed tallest first into resolution^2 with
//    `padding` texels between charts, at the largest density that fits
//    (bisection); lightmap UVs are texel positions / resolution;
//  * output: vertices split per chart (xref to the source vertex), faces in
//    their original order, transform / material / other attributes kept.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/shb.h"

struct shb_atlas {
  std::vector<float> pos, nrm, tuv, luv, tan;
  std::vector<uint32_t> idx, xref, chart_of;
  std::vector<shb_geometry> geos;
  std::vector<shb_material> mats;
  std::vector<uint8_t> texels;
  std::vector<shb_light> lights;
  shb_scene_desc desc{};
  int32_t width = 0, height = 0;
  float density = 0.f;
  uint32_t charts = 0;
};

namespace {

constexpr float kMinScaleRange = 0.001f;  // atlas.cpp:59-61
constexpr float kMaxScaleFactor = 5.0f;   // atlas.cpp:16
constexpr double kChartCos = 0.9063077870366499;  // cos 25 degrees

struct D3 {
  double x, y, z;
};
D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
double norm(D3 a) { return std::sqrt(dot(a, a)); }

struct P2 {
  double x, y;
};

double cross2(P2 o, P2 a, P2 b) { return (a.x - o.x) * (b.y - o.y) - (a.y - o.y) * (b.x - o.x); }

// Andrew's monotone chain
std::vector<P2> hull(std::vector<P2> p) {
  std::sort(p.begin(), p.end(), [](P2 a, P2 b) { return a.x < b.x || (a.x == b.x && a.y < b.y); });
  p.erase(std::unique(p.begin(), p.end(), [](P2 a, P2 b) { return a.x == b.x && a.y == b.y; }),
          p.end());
  if (p.size() < 3) return p;
  std::vector<P2> h(2 * p.size());
  size_t k = 0;
  for (size_t i = 0; i < p.size(); ++i) {
    while (k >= 2 && cross2(h[k - 2], h[k - 1], p[i]) <= 0) --k;
    h[k++] = p[i];
  }
  for (size_t i = p.size() - 1, t = k + 1; i-- > 0;) {
    while (k >= t && cross2(h[k - 2], h[k - 1], p[i]) <= 0) --k;
    h[k++] = p[i];
  }
  h.resize(k - 1);
  return h;
}

struct Chart {
  std::vector<uint32_t> faces;  // global face ids within the geometry
  D3 axis{0, 0, 1};
  D3 e1{1, 0, 0}, e2{0, 1, 0};  // in-plane basis after the calipers rotation
  double u0 = 0, v0 = 0, w = 0, h = 0;  // bbox in (e1, e2)
  double px = 0, py = 0;  // packed origin (texels)
  int geo = 0;
};

void put_err(const std::string& e, char* err, size_t len) {
  if (!err || !len) return;
  std::strncpy(err, e.c_str(), len - 1);
  err[len - 1] = 0;
}

bool pack(std::vector<Chart>& charts, double d, int res, int pad, bool commit) {
  std::vector<size_t> order(charts.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return charts[a].h > charts[b].h; });
  long long x = 0, y = 0, row_h = 0;
  for (size_t ci : order) {
    const long long rw = (long long)std::ceil(charts[ci].w * d) + 1 + pad;
    const long long rh = (long long)std::ceil(charts[ci].h * d) + 1 + pad;
    if (rw > res) return false;
    if (x + rw > res) {
      x = 0;
      y += row_h;
      row_h = 0;
    }
    if (y + rh > res) return false;
    if (commit) {
      charts[ci].px = (double)x + 0.5 * (pad + 1);
      charts[ci].py = (double)y + 0.5 * (pad + 1);
    }
    x += rw;
    row_h = std::max(row_h, rh);
  }
  return true;
}

}  // namespace

extern "C" {

// src/atlas.cpp:22-95 CalculateGeometryScales, float arithmetic as there
int shb_atlas_geometry_scales(const shb_scene_desc* s, float density_multiplier, float* out) {
  if (!s || !out) return SHB_ERR_INVALID_ARG;
  std::vector<float> scales(s->num_geometries);
  for (uint32_t gi = 0; gi < s->num_geometries; ++gi) {
    const shb_geometry& g = s->geometries[gi];
    if (g.material_id < 0 || (uint32_t)g.material_id >= s->num_materials) return SHB_ERR_INVALID_ARG;
    if (!(g.flags & SHB_GEOM_HAS_TEXTURE_UVS) || !s->texture_uvs || g.vertex_count == 0)
      return SHB_ERR_INVALID_ARG;
    const shb_material& m = s->materials[g.material_id];
    const float width = (float)m.albedo_width, height = (float)m.albedo_height;
    const float tex_area_sqrt = std::sqrt(width * height);
    const float target_scale = density_multiplier * tex_area_sqrt;
    float u_min = 3.402823466e38f, u_max = -3.402823466e38f;
    float v_min = 3.402823466e38f, v_max = -3.402823466e38f;
    for (uint32_t v = g.vertex_offset; v < g.vertex_offset + g.vertex_count; ++v) {
      const float u = s->texture_uvs[2 * v], w = s->texture_uvs[2 * v + 1];
      u_min = std::min(u_min, u);
      u_max = std::max(u_max, u);
      v_min = std::min(v_min, w);
      v_max = std::max(v_max, w);
    }
    const float u_range = std::max(u_max - u_min, kMinScaleRange);
    const float v_range = std::max(v_max - v_min, kMinScaleRange);
    const float tile_count = u_range * v_range;
    scales[gi] = target_scale * std::sqrt(tile_count);
  }
  if (!scales.empty()) {
    std::vector<float> sorted = scales;
    std::sort(sorted.begin(), sorted.end());
    const float median = std::max(sorted[sorted.size() / 2], 0.001f);
    const float max_allowed = kMaxScaleFactor * median;
    for (float& v : scales) v = std::min(v, max_allowed);
  }
  std::copy(scales.begin(), scales.end(), out);
  return SHB_OK;
}

int shb_atlas_create(const shb_scene_desc* s, uint32_t resolution, uint32_t padding,
                     float density_multiplier, shb_atlas** out, char* err, size_t err_len) {
  if (!s || !out || resolution == 0 || resolution > (1u << 16) || padding > 64) {
    put_err("bad arguments", err, err_len);
    return SHB_ERR_INVALID_ARG;
  }
  *out = nullptr;
  if (s->num_geometries == 0) {
    put_err("no geometries", err, err_len);
    return SHB_ERR_INVALID_ARG;
  }
  if (!s->normals || !s->tangents) {
    put_err("the atlas needs normals and tangents on every geometry", err, err_len);
    return SHB_ERR_INVALID_ARG;
  }
  std::vector<float> scales(s->num_geometries);
  if (shb_atlas_geometry_scales(s, density_multiplier, scales.data()) != SHB_OK) {
    put_err("every geometry needs a valid material and texture UVs", err, err_len);
    return SHB_ERR_INVALID_ARG;
  }
  auto* a = new shb_atlas();
  std::vector<Chart> charts;
  // per geometry: world positions (scaled), faces, charts
  std::vector<std::vector<uint32_t>> face_chart(s->num_geometries);
  std::vector<std::vector<P2>> corner_uv(s->num_geometries);  // 3 per face, chart plane coords
  for (uint32_t gi = 0; gi < s->num_geometries; ++gi) {
    const shb_geometry& g = s->geometries[gi];
    if ((uint64_t)g.vertex_offset + g.vertex_count > s->num_vertices ||
        (uint64_t)g.index_offset + g.index_count > s->num_indices || g.index_count % 3) {
      delete a;
      put_err("geometry ranges out of bounds", err, err_len);
      return SHB_ERR_INVALID_ARG;
    }
    const float* T = g.transform;
    const uint32_t nv = g.vertex_count, nf = g.index_count / 3;
    std::vector<D3> P(nv);
    for (uint32_t v = 0; v < nv; ++v) {
      const float* p = &s->positions[3 * (size_t)(g.vertex_offset + v)];
      // TransformedVertices (scene.cpp:162-169) in float, then the scale
      const float x = T[0] * p[0] + T[1] * p[1] + T[2] * p[2] + T[3];
      const float y = T[4] * p[0] + T[5] * p[1] + T[6] * p[2] + T[7];
      const float z = T[8] * p[0] + T[9] * p[1] + T[10] * p[2] + T[11];
      P[v] = {(double)(x * scales[gi]), (double)(y * scales[gi]), (double)(z * scales[gi])};
    }
    const uint32_t* I = &s->indices[g.index_offset];
    for (uint32_t k = 0; k < g.index_count; ++k)
      if (I[k] >= nv) {
        delete a;
        put_err("index out of range", err, err_len);
        return SHB_ERR_INVALID_ARG;
      }
    // welded position ids
    std::vector<uint32_t> pid(nv);
    {
      struct H {
        size_t operator()(const D3& d) const {
          size_t h = std::hash<double>()(d.x);
          h = h * 1000003u ^ std::hash<double>()(d.y);
          return h * 1000003u ^ std::hash<double>()(d.z);
        }
      };
      struct E {
        bool operator()(const D3& a, const D3& b) const { return a.x == b.x && a.y == b.y && a.z == b.z; }
      };
      std::unordered_map<D3, uint32_t, H, E> m;
      m.reserve(nv * 2);
      for (uint32_t v = 0; v < nv; ++v) pid[v] = m.emplace(P[v], (uint32_t)m.size()).first->second;
    }
    std::vector<D3> fn(nf);
    std::vector<double> farea(nf);
    for (uint32_t f = 0; f < nf; ++f) {
      const D3 c = cross(sub(P[I[3 * f + 1]], P[I[3 * f]]), sub(P[I[3 * f + 2]], P[I[3 * f]]));
      const double l = norm(c);
      farea[f] = 0.5 * l;
      fn[f] = l > 0 ? D3{c.x / l, c.y / l, c.z / l} : D3{0, 0, 0};
    }
    // edge -> faces
    std::unordered_map<uint64_t, std::vector<uint32_t>> edges;
    edges.reserve(nf * 3);
    for (uint32_t f = 0; f < nf; ++f)
      for (int e = 0; e < 3; ++e) {
        uint32_t p0 = pid[I[3 * f + e]], p1 = pid[I[3 * f + (e + 1) % 3]];
        if (p0 > p1) std::swap(p0, p1);
        if (p0 != p1) edges[(uint64_t)p0 << 32 | p1].push_back(f);
      }
    std::vector<uint32_t> order(nf);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return farea[x] > farea[y]; });
    std::vector<int> assigned(nf, -1);
    std::vector<uint32_t> queue;
    face_chart[gi].assign(nf, 0);
    corner_uv[gi].assign((size_t)nf * 3, P2{0, 0});
    for (uint32_t seed : order) {
      if (assigned[seed] >= 0) continue;
      Chart ch;
      ch.geo = (int)gi;
      const int cid = (int)charts.size();
      const D3 sn = fn[seed];
      D3 acc = {sn.x * farea[seed], sn.y * farea[seed], sn.z * farea[seed]};
      D3 axis = sn;
      assigned[seed] = cid;
      queue.assign(1, seed);
      const bool degenerate = farea[seed] == 0.0;
      for (size_t qi = 0; qi < queue.size() && !degenerate; ++qi) {
        const uint32_t f = queue[qi];
        for (int e = 0; e < 3; ++e) {
          uint32_t p0 = pid[I[3 * f + e]], p1 = pid[I[3 * f + (e + 1) % 3]];
          if (p0 > p1) std::swap(p0, p1);
          auto it = edges.find((uint64_t)p0 << 32 | p1);
          if (it == edges.end()) continue;
          for (uint32_t nb : it->second) {
            if (assigned[nb] >= 0 || farea[nb] == 0.0) continue;
            if (dot(fn[nb], sn) < kChartCos || dot(fn[nb], axis) < kChartCos) continue;
            assigned[nb] = cid;
            queue.push_back(nb);
            acc = {acc.x + fn[nb].x * farea[nb], acc.y + fn[nb].y * farea[nb], acc.z + fn[nb].z * farea[nb]};
            const double l = norm(acc);
            if (l > 0) axis = {acc.x / l, acc.y / l, acc.z / l};
          }
        }
      }
      ch.faces = queue;
      // project along the final mean normal only if every face stays within
      // 25 degrees of it (it drifted while growing); else along the seed
      // normal, which every face is within 25 degrees of
      for (uint32_t f : ch.faces)
        if (dot(fn[f], axis) < kChartCos) {
          axis = sn;
          break;
        }
      if (norm(axis) == 0) axis = {0, 0, 1};
      ch.axis = axis;
      // in-plane basis
      D3 ref = std::fabs(axis.x) < 0.9 ? D3{1, 0, 0} : D3{0, 1, 0};
      D3 b1 = cross(ref, axis);
      b1 = {b1.x / norm(b1), b1.y / norm(b1), b1.z / norm(b1)};
      const D3 b2 = cross(axis, b1);
      std::vector<P2> pts;
      pts.reserve(ch.faces.size() * 3);
      for (uint32_t f : ch.faces)
        for (int k = 0; k < 3; ++k) {
          const D3& q = P[I[3 * f + k]];
          pts.push_back({dot(q, b1), dot(q, b2)});
        }
      // minimum-area rectangle over hull edge directions
      std::vector<P2> hp = hull(pts);
      double best = 1e300, bc = 1, bs = 0;
      const size_t hn = hp.size();
      const size_t step = std::max<size_t>(1, hn / 256);
      for (size_t i = 0; i < std::max<size_t>(hn, 1); i += step) {
        double cx = 1, sy = 0;
        if (hn >= 2) {
          const P2 d = {hp[(i + 1) % hn].x - hp[i].x, hp[(i + 1) % hn].y - hp[i].y};
          const double l = std::sqrt(d.x * d.x + d.y * d.y);
          if (l == 0) continue;
          cx = d.x / l;
          sy = d.y / l;
        }
        double umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300;
        for (const P2& q : hp) {
          const double u = q.x * cx + q.y * sy, v = -q.x * sy + q.y * cx;
          umin = std::min(umin, u);
          umax = std::max(umax, u);
          vmin = std::min(vmin, v);
          vmax = std::max(vmax, v);
        }
        const double area = (umax - umin) * (vmax - vmin);
        if (area < best) {
          best = area;
          bc = cx;
          bs = sy;
        }
      }
      // e1 = b1 cos + b2 sin, e2 = -b1 sin + b2 cos; put the long side along u
      D3 e1 = {b1.x * bc + b2.x * bs, b1.y * bc + b2.y * bs, b1.z * bc + b2.z * bs};
      D3 e2 = {-b1.x * bs + b2.x * bc, -b1.y * bs + b2.y * bc, -b1.z * bs + b2.z * bc};
      double umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300;
      for (uint32_t f : ch.faces)
        for (int k = 0; k < 3; ++k) {
          const D3& q = P[I[3 * f + k]];
          const double u = dot(q, e1), v = dot(q, e2);
          umin = std::min(umin, u);
          umax = std::max(umax, u);
          vmin = std::min(vmin, v);
          vmax = std::max(vmax, v);
        }
      ch.e1 = e1;
      ch.e2 = e2;
      ch.u0 = umin;
      ch.v0 = vmin;
      ch.w = std::max(0.0, umax - umin);
      ch.h = std::max(0.0, vmax - vmin);
      for (uint32_t f : ch.faces) {
        face_chart[gi][f] = (uint32_t)cid;
        for (int k = 0; k < 3; ++k) {
          const D3& q = P[I[3 * f + k]];
          corner_uv[gi][(size_t)f * 3 + k] = {dot(q, e1) - umin, dot(q, e2) - vmin};
        }
      }
      charts.push_back(std::move(ch));
    }
  }
  // densest packing that fits: bisection on texels per (scaled) unit
  const int res = (int)resolution, pad = (int)padding;
  double area = 0, maxside = 0;
  for (const Chart& c : charts) {
    area += (c.w + 1e-12) * (c.h + 1e-12);
    maxside = std::max(maxside, std::max(c.w, c.h));
  }
  double lo = 0.0, hi = maxside > 0 ? (res - 1 - pad) / maxside : 1.0;
  if (area > 0) hi = std::min(hi, 2.0 * res / std::sqrt(area));
  if (!pack(charts, lo, res, pad, false)) {
    delete a;
    put_err("the charts do not fit a " + std::to_string(res) + "^2 atlas with padding " +
                std::to_string(pad) + " (" + std::to_string(charts.size()) + " charts)",
            err, err_len);
    return SHB_ERR_LIMIT;
  }
  while (pack(charts, hi, res, pad, false) && hi < 1e30) {
    lo = hi;
    hi *= 2;
  }
  for (int it = 0; it < 40; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (pack(charts, mid, res, pad, false)) lo = mid;
    else hi = mid;
  }
  const double d = lo;
  pack(charts, d, res, pad, true);
  // output: vertices split per (chart, source vertex)
  for (uint32_t gi = 0; gi < s->num_geometries; ++gi) {
    const shb_geometry& g = s->geometries[gi];
    const uint32_t nf = g.index_count / 3;
    const uint32_t* I = &s->indices[g.index_offset];
    shb_geometry ng = g;
    ng.vertex_offset = (uint32_t)(a->pos.size() / 3);
    ng.index_offset = (uint32_t)a->idx.size();
    ng.flags |= SHB_GEOM_HAS_LIGHTMAP_UVS;
    std::unordered_map<uint64_t, uint32_t> remap;
    remap.reserve(nf * 3);
    uint32_t count = 0;
    for (uint32_t f = 0; f < nf; ++f) {
      const uint32_t cid = face_chart[gi][f];
      const Chart& ch = charts[cid];
      for (int k = 0; k < 3; ++k) {
        const uint32_t v = I[3 * f + k];
        auto ins = remap.emplace((uint64_t)cid << 32 | v, count);
        if (ins.second) {
          const size_t sv = (size_t)g.vertex_offset + v;
          a->pos.insert(a->pos.end(), &s->positions[3 * sv], &s->positions[3 * sv] + 3);
          a->nrm.insert(a->nrm.end(), &s->normals[3 * sv], &s->normals[3 * sv] + 3);
          if (s->texture_uvs)
            a->tuv.insert(a->tuv.end(), &s->texture_uvs[2 * sv], &s->texture_uvs[2 * sv] + 2);
          else
            a->tuv.insert(a->tuv.end(), 2, 0.0f);
          a->tan.insert(a->tan.end(), &s->tangents[4 * sv], &s->tangents[4 * sv] + 4);
          const P2 q = corner_uv[gi][(size_t)f * 3 + k];
          a->luv.push_back((float)((ch.px + q.x * d) / res));
          a->luv.push_back((float)((ch.py + q.y * d) / res));
          a->xref.push_back(v);
          a->chart_of.push_back(cid);
          ++count;
        }
        a->idx.push_back(ins.first->second);  // geometry-local, as in the input
      }
    }
    ng.vertex_count = count;
    ng.index_count = nf * 3;
    a->geos.push_back(ng);
  }
  a->mats.assign(s->materials, s->materials + s->num_materials);
  a->texels.assign(s->texel_pool, s->texel_pool + s->texel_pool_size);
  a->lights.assign(s->lights, s->lights + s->num_lights);
  a->width = a->height = res;
  a->density = (float)d;
  a->charts = (uint32_t)charts.size();
  shb_scene_desc& o = a->desc;
  o.num_vertices = (uint32_t)(a->pos.size() / 3);
  o.positions = a->pos.data();
  o.normals = a->nrm.data();
  o.texture_uvs = a->tuv.data();
  o.lightmap_uvs = a->luv.data();
  o.tangents = a->tan.data();
  o.num_indices = (uint32_t)a->idx.size();
  o.indices = a->idx.data();
  o.num_geometries = (uint32_t)a->geos.size();
  o.geometries = a->geos.data();
  o.num_materials = (uint32_t)a->mats.size();
  o.materials = a->mats.data();
  o.texel_pool_size = a->texels.size();
  o.texel_pool = a->texels.data();
  o.num_lights = (uint32_t)a->lights.size();
  o.lights = a->lights.data();
  *out = a;
  return SHB_OK;
}

const shb_scene_desc* shb_atlas_scene(const shb_atlas* a) { return a ? &a->desc : nullptr; }

int shb_atlas_info(const shb_atlas* a, int32_t* width, int32_t* height, uint32_t* charts,
                   float* texels_per_unit, const uint32_t** xref, const uint32_t** chart) {
  if (!a) return SHB_ERR_INVALID_ARG;
  if (chart) *chart = a->chart_of.data();
  if (width) *width = a->width;
  if (height) *height = a->height;
  if (charts) *charts = a->charts;
  if (texels_per_unit) *texels_per_unit = a->density;
  if (xref) *xref = a->xref.data();
  return SHB_OK;
}

void shb_atlas_free(shb_atlas* a) { delete a; }

}  // extern "C"
