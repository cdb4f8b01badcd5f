// Host-side scene ingestion (see scene_host.h). Compiled with
// -ffp-contract=off: world transforms and area-light pdfs round exactly as
// the kernels and the host restatement expect.
#include "scene_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace shb {
namespace {

// ------------------------------------------------------ host float math
// Same evaluation order as the kernels (e0 + (e1 + e2) reductions); this
// file is compiled with -ffp-contract=off.
struct H3 {
  float x, y, z;
};
inline H3 h3sub(H3 a, H3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float h3sq(H3 a) { return a.x * a.x + (a.y * a.y + a.z * a.z); }
inline H3 h3norm(H3 a) {
  float z = h3sq(a);
  if (z > 0.f) {
    float s = std::sqrt(z);
    return {a.x / s, a.y / s, a.z / s};
  }
  return a;
}
inline H3 h3cross(H3 a, H3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float rowdot(const float* r, H3 v) { return r[0] * v.x + (r[1] * v.y + r[2] * v.z); }

// Unitary factor of the polar decomposition of the 3x3 part (what Eigen's
// Transform::rotation() returns), by Newton iteration R <- (R + R^-T)/2.
void rotation_of(const float* xf, float R[9]) {
  double m[9] = {xf[0], xf[1], xf[2], xf[4], xf[5], xf[6], xf[8], xf[9], xf[10]};
  for (int it = 0; it < 64; ++it) {
    double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7],
           i = m[8];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    double det = a * A + b * B + c * C;
    if (det == 0.0) break;
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

}  // namespace

// src/colorspace.h:10-14
float srgb_to_linear(uint8_t x) {
  float v = x / 255.0f;
  if (v <= 0.04045f) return v / 12.92f;
  return std::pow((v + 0.055f) / 1.055f, 2.4f);
}


int build_host_scene(const shb_scene_desc* d, HostScene* out, std::string* err) {
  HostScene hs;
  auto fail = [&](int code, const std::string& m) {
    *err = m;
    return code;
  };
  if (!d) return fail(SHB_ERR_INVALID_ARG, "scene is NULL");
  if (d->num_vertices && !d->positions) return fail(SHB_ERR_INVALID_ARG, "positions is NULL");
  if (d->num_indices && !d->indices) return fail(SHB_ERR_INVALID_ARG, "indices is NULL");
  if (d->num_geometries && !d->geometries)
    return fail(SHB_ERR_INVALID_ARG, "geometries is NULL");
  if (d->num_materials && !d->materials) return fail(SHB_ERR_INVALID_ARG, "materials is NULL");
  if (d->num_lights && !d->lights) return fail(SHB_ERR_INVALID_ARG, "lights is NULL");
  hs.num_vertices = d->num_vertices;
  // materials (src/scene.h:43-53)
  for (uint32_t m = 0; m < d->num_materials; ++m) {
    const shb_material& sm = d->materials[m];
    MaterialDev md{};
    md.emission = sm.emission_intensity;
    if (sm.albedo_channels != 0) {
      if (sm.albedo_channels < 3 || sm.albedo_channels > 4 || sm.albedo_width == 0 ||
          sm.albedo_height == 0)
        return fail(SHB_ERR_INVALID_ARG, "material " + std::to_string(m) + ": bad texture");
      uint64_t nb = (uint64_t)sm.albedo_width * sm.albedo_height * sm.albedo_channels;
      if (!d->texel_pool || sm.albedo_offset + nb > d->texel_pool_size)
        return fail(SHB_ERR_INVALID_ARG,
                    "material " + std::to_string(m) + ": texture outside the texel pool");
      md.tex_w = sm.albedo_width;
      md.tex_h = sm.albedo_height;
      md.tex_c = sm.albedo_channels;
      md.tex_off = sm.albedo_offset;
      if (sm.albedo_channels == 4) {
        const uint8_t* px = d->texel_pool + sm.albedo_offset;
        for (uint64_t k = 3; k < nb; k += 4)
          if (px[k] < 255) hs.has_transparency = true;
      }
    }
    hs.mats.push_back(md);
  }
  if (d->texel_pool_size) hs.texels.assign(d->texel_pool, d->texel_pool + d->texel_pool_size);
  // geometries -> global triangles
  std::vector<int> geom_first_tri(d->num_geometries, 0);
  hs.wpos.assign(3 * (size_t)d->num_vertices, 0.f);
  hs.wnrm.assign(3 * (size_t)d->num_vertices, 0.f);
  hs.wtan.assign(4 * (size_t)d->num_vertices, 0.f);
  hs.luv.assign(2 * (size_t)d->num_vertices, 0.f);
  for (uint32_t gi = 0; gi < d->num_geometries; ++gi) {
    const shb_geometry& g = d->geometries[gi];
    std::string gs = "geometry " + std::to_string(gi);
    if ((uint64_t)g.vertex_offset + g.vertex_count > d->num_vertices)
      return fail(SHB_ERR_INVALID_ARG, gs + ": vertices out of range");
    if ((uint64_t)g.index_offset + g.index_count > d->num_indices)
      return fail(SHB_ERR_INVALID_ARG, gs + ": indices out of range");
    // material ids only matter to the bake (baker.cpp:94 indexes materials
    // with them); rasterization and ray queries pass them through as the
    // reference does (rasterizer.cpp:133-151, occlusion.cpp:50).
    if (g.index_count > 0 && (g.material_id < 0 || (uint32_t)g.material_id >= d->num_materials))
      hs.bad_material = true;
    bool hasN = (g.flags & SHB_GEOM_HAS_NORMALS) && d->normals;
    bool hasUV = (g.flags & SHB_GEOM_HAS_TEXTURE_UVS) && d->texture_uvs;
    bool hasL = (g.flags & SHB_GEOM_HAS_LIGHTMAP_UVS) && d->lightmap_uvs;
    bool hasT = (g.flags & SHB_GEOM_HAS_TANGENTS) && d->tangents;
    if (!hasL || !hasN || !hasT) hs.raster_ready = false;
    float R[9];
    rotation_of(g.transform, R);
    // world attributes (src/scene.cpp:162-195)
    for (uint32_t k = 0; k < g.vertex_count; ++k) {
      uint32_t vi = g.vertex_offset + k;
      H3 p = {d->positions[3 * vi], d->positions[3 * vi + 1], d->positions[3 * vi + 2]};
      hs.wpos[3 * vi] = rowdot(g.transform, p) + g.transform[3];
      hs.wpos[3 * vi + 1] = rowdot(g.transform + 4, p) + g.transform[7];
      hs.wpos[3 * vi + 2] = rowdot(g.transform + 8, p) + g.transform[11];
      if (hasN) {
        H3 n = {d->normals[3 * vi], d->normals[3 * vi + 1], d->normals[3 * vi + 2]};
        H3 wn = h3norm(H3{rowdot(R, n), rowdot(R + 3, n), rowdot(R + 6, n)});
        hs.wnrm[3 * vi] = wn.x;
        hs.wnrm[3 * vi + 1] = wn.y;
        hs.wnrm[3 * vi + 2] = wn.z;
      }
      if (hasT) {
        H3 t = {d->tangents[4 * vi], d->tangents[4 * vi + 1], d->tangents[4 * vi + 2]};
        H3 wt = h3norm(H3{rowdot(R, t), rowdot(R + 3, t), rowdot(R + 6, t)});
        hs.wtan[4 * vi] = wt.x;
        hs.wtan[4 * vi + 1] = wt.y;
        hs.wtan[4 * vi + 2] = wt.z;
        hs.wtan[4 * vi + 3] = d->tangents[4 * vi + 3];
      }
      if (hasL) {
        hs.luv[2 * vi] = d->lightmap_uvs[2 * vi];
        hs.luv[2 * vi + 1] = d->lightmap_uvs[2 * vi + 1];
      }
    }
    uint32_t nt = g.index_count / 3;
    geom_first_tri[gi] = (int)hs.num_tris;
    for (uint32_t t = 0; t < nt; ++t) {
      uint32_t li[3];
      for (int k = 0; k < 3; ++k) {
        li[k] = d->indices[g.index_offset + 3 * t + k];
        if (li[k] >= g.vertex_count)
          return fail(SHB_ERR_INVALID_ARG, gs + ": index out of range");
      }
      uint32_t gv[3] = {g.vertex_offset + li[0], g.vertex_offset + li[1], g.vertex_offset + li[2]};
      H3 w[3];
      for (int k = 0; k < 3; ++k) {
        w[k] = {hs.wpos[3 * gv[k]], hs.wpos[3 * gv[k] + 1], hs.wpos[3 * gv[k] + 2]};
        hs.world_xyz.push_back(w[k].x);
        hs.world_xyz.push_back(w[k].y);
        hs.world_xyz.push_back(w[k].z);
        hs.tri_vidx.push_back(gv[k]);
      }
      hs.tri_mat.push_back(g.material_id);
      // shading record (src/occlusion.cpp:63-93: object-space normals)
      float n[9] = {0}, uv[6] = {0};
      uint32_t flags = 0;
      if (hasN) {
        flags |= 1u;
        for (int k = 0; k < 3; ++k)
          for (int a = 0; a < 3; ++a) n[3 * k + a] = d->normals[3 * gv[k] + a];
      } else {
        H3 ng = h3norm(h3cross(h3sub(w[1], w[0]), h3sub(w[2], w[0])));
        n[0] = ng.x;
        n[1] = ng.y;
        n[2] = ng.z;
      }
      if (hasUV) {
        flags |= 2u;
        for (int k = 0; k < 3; ++k) {
          uv[2 * k] = d->texture_uvs[2 * gv[k]];
          uv[2 * k + 1] = d->texture_uvs[2 * gv[k] + 1];
        }
      }
      hs.shade.push_back(make_float4(n[0], n[1], n[2], uv[0]));
      hs.shade.push_back(make_float4(n[3], n[4], n[5], uv[1]));
      hs.shade.push_back(make_float4(n[6], n[7], n[8], uv[2]));
      hs.shade.push_back(make_float4(uv[3], uv[4], uv[5], as_float(flags)));
      hs.tri_info.push_back(make_int4(g.material_id, -1, (int)gi, (int)t));
      hs.num_tris++;
    }
  }
  // lights (src/scene.h:70-87, src/light.h, src/light.cpp:16-78)
  for (uint32_t li = 0; li < d->num_lights; ++li) {
    const shb_light& sl = d->lights[li];
    LightDev L{};
    L.type = sl.type;
    if (sl.type < SHB_LIGHT_POINT || sl.type > SHB_LIGHT_AREA)
      return fail(SHB_ERR_INVALID_ARG, "light " + std::to_string(li) + ": bad type");
    for (int a = 0; a < 3; ++a) {
      L.ic[a] = sl.intensity * sl.color[a];
      L.pos[a] = sl.position[a];
      L.dir[a] = sl.direction[a];
    }
    L.cos_inner = sl.cos_inner_cone;
    L.cos_outer = sl.cos_outer_cone;
    L.material = -1;
    if (sl.type == SHB_LIGHT_AREA) {
      int gi = sl.geometry_index, mi = sl.material_index;
      if (gi >= 0 && (uint32_t)gi < d->num_geometries && mi >= 0 &&
          (uint32_t)mi < d->num_materials && d->geometries[gi].index_count >= 3) {
        const shb_geometry& g = d->geometries[gi];
        uint32_t nt = g.index_count / 3;
        L.first_tri = (uint32_t)hs.area_tris.size();
        L.num_tris = nt;
        L.material = mi;
        bool hasUV = (g.flags & SHB_GEOM_HAS_TEXTURE_UVS) && d->texture_uvs;
        for (uint32_t t = 0; t < nt; ++t) {
          AreaTri at{};
          H3 v[3];
          for (int k = 0; k < 3; ++k) {
            uint32_t vi = g.vertex_offset + d->indices[g.index_offset + 3 * t + k];
            v[k] = {d->positions[3 * vi], d->positions[3 * vi + 1], d->positions[3 * vi + 2]};
            if (hasUV) {
              at.uv[2 * k] = d->texture_uvs[2 * vi];
              at.uv[2 * k + 1] = d->texture_uvs[2 * vi + 1];
            }
          }
          at.v0[0] = v[0].x, at.v0[1] = v[0].y, at.v0[2] = v[0].z;
          at.e1[0] = v[1].x - v[0].x, at.e1[1] = v[1].y - v[0].y, at.e1[2] = v[1].z - v[0].z;
          at.e2[0] = v[2].x - v[0].x, at.e2[1] = v[2].y - v[0].y, at.e2[2] = v[2].z - v[0].z;
          float area = std::sqrt(h3sq(h3cross(h3sub(v[0], v[1]), h3sub(v[0], v[2])))) / 2.f;
          at.pdf = std::max(1e-6f, 1.f / (float)nt * 1.f / area);
          at.has_uv = hasUV ? 1.f : 0.f;
          at.leaf = kNoLeaf;  // set from the BVH at upload
          hs.area_tris.push_back(at);
          hs.area_tri_global.push_back((uint32_t)(geom_first_tri[gi] + (int)t));
        }
        const MaterialDev& md = hs.mats[mi];
        if (md.tex_c == 0 || (md.tex_w == 1 && md.tex_h == 1)) {
          float alb[3] = {0.8f, 0.8f, 0.8f};
          if (md.tex_c != 0)
            for (int a = 0; a < 3; ++a) alb[a] = srgb_to_linear(hs.texels[md.tex_off + a]);
          L.emit_const = 1;
          for (int a = 0; a < 3; ++a)
            L.emit[a] = md.emission > 0.0f ? alb[a] * md.emission : 0.f;
        }
        // triangles of the emitting geometry carry the light for the
        // emission they show when hit directly
        for (uint32_t t = 0; t < nt; ++t) hs.tri_info[geom_first_tri[gi] + t].y = (int)li;
      }
    }
    // the channels the selection weight scales: intensity x colour, or a
    // constant emitter's emission (a textured emitter's vary per sample)
    const float* wc = (sl.type == SHB_LIGHT_AREA) ? L.emit : L.ic;
    const bool per_sample = sl.type == SHB_LIGHT_AREA && !L.emit_const;
    L.wmax = per_sample ? 0.f : std::max(wc[0], std::max(wc[1], wc[2]));
    L.grey = !per_sample && wc[0] == wc[1] && wc[1] == wc[2] ? 1 : 0;
    hs.lights.push_back(L);
  }
  // runs of simple area lights, counted from the back
  for (size_t i = hs.lights.size(); i-- > 0;) {
    LightDev& L = hs.lights[i];
    const bool simple = L.type == kLightArea && L.num_tris == 1 && L.emit_const && L.grey;
    const uint32_t next = i + 1 < hs.lights.size() ? hs.lights[i + 1].simple_run : 0u;
    L.simple_run = simple ? next + 1u : 0u;
  }
  *out = std::move(hs);
  return SHB_OK;
}


}  // namespace shb
