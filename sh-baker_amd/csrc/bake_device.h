// Device-side shading of the bake: materials, lights, SH9, sampling.
// Restates the reference's per-vertex arithmetic in the same operation order
// (compiled with -ffp-contract=off); references are to DaviesX/sh-baker.
#ifndef SHB_BAKE_DEVICE_H_
#define SHB_BAKE_DEVICE_H_

#include "shb_common.h"
#include "traverse.h"

namespace shb {

constexpr double kPi = 3.14159265358979323846;

// src/material.cpp:69-85 GetAlbedo: nearest texel, clamp, per-fetch sRGB
// decode (256-entry table equal to SRGBToLinear, colorspace.h:10-14).
__device__ __forceinline__ f3 get_albedo(const SceneView& sc, int m, float u, float v) {
  const MaterialDev& M = sc.mats[m];
  if (M.tex_c == 0) return {0.8f, 0.8f, 0.8f};
  int tx = min(max((int)(u * (float)M.tex_w), 0), (int)M.tex_w - 1);
  int ty = min(max((int)(v * (float)M.tex_h), 0), (int)M.tex_h - 1);
  const uint8_t* px = sc.texels + M.tex_off + ((size_t)ty * M.tex_w + tx) * M.tex_c;
  return {sc.srgb_lut[px[0]], sc.srgb_lut[px[1]], sc.srgb_lut[px[2]]};
}

// src/material.cpp:57-67 GetAlpha.
__device__ __forceinline__ float get_alpha(const SceneView& sc, int m, float u, float v) {
  const MaterialDev& M = sc.mats[m];
  if (M.tex_c < 4) return 1.0f;
  int tx = min(max((int)(u * (float)M.tex_w), 0), (int)M.tex_w - 1);
  int ty = min(max((int)(v * (float)M.tex_h), 0), (int)M.tex_h - 1);
  const uint8_t* px = sc.texels + M.tex_off + ((size_t)ty * M.tex_w + tx) * M.tex_c;
  return px[3] / 255.0f;
}

// src/material.cpp:87-92 GetEmission (unscaled albedo x intensity).
__device__ __forceinline__ f3 get_emission(const SceneView& sc, int m, float u, float v) {
  float e = sc.mats[m].emission;
  if (e > 0.0f) return get_albedo(sc, m, u, v) * e;
  return {0.f, 0.f, 0.f};
}

// Lambert BRDF value albedo/pi (material.cpp:160-161) with the material's
// albedo scale (1 reproduces the reference); `unscaled` receives the value
// without the scale.
__device__ __forceinline__ f3 lambert(const SceneView& sc, int m, float u, float v,
                                      f3* unscaled = nullptr) {
  f3 a = get_albedo(sc, m, u, v);
  if (unscaled) *unscaled = a / (float)kPi;
  const float* s = sc.albedo_scale + 3 * m;
  a = cmul(a, f3{s[0], s[1], s[2]});
  return a / (float)kPi;
}

// src/baker.cpp:38-46 SampleHemisphereUniform given its uniforms.
__device__ __forceinline__ f3 hemisphere(float u1, float u2) {
  float r = sqrtf(1.0f - u1 * u1);
  float phi = (float)(2.0 * kPi * (double)u2);
  return {r * cosf(phi), r * sinf(phi), u1};
}

// src/material.cpp:113-146 SampleMaterial (Lambert, cosine-weighted).
__device__ __forceinline__ void sample_lambert(f3 n, float r1, float r2, f3& dir, float& pdf) {
  f3 t;
  if (fabsf(n.x) > fabsf(n.z))
    t = {-n.y, n.x, 0.0f};
  else
    t = {0.0f, -n.z, n.y};
  t = normalized(t);
  f3 b = cross(n, t);
  float phi = (float)(2.0 * kPi * (double)r1);
  float theta = acosf(sqrtf(r2));
  float sin_theta = sinf(theta);
  f3 l = {cosf(phi) * sin_theta, sinf(phi) * sin_theta, cosf(theta)};
  f3 w = t * l.x + b * l.y + n * l.z;
  dir = normalized(w);
  pdf = (float)((double)l.z / kPi);
}

// src/sh_coeffs.cpp:50-78 basis (direction not normalised, constants :27-31).
__device__ __forceinline__ void sh_basis(f3 d, float Y[9]) {
  const float x = d.x, y = d.y, z = d.z;
  Y[0] = 0.282095f;
  Y[1] = 0.488603f * y;
  Y[2] = 0.488603f * z;
  Y[3] = 0.488603f * x;
  Y[4] = 1.092548f * x * y;
  Y[5] = 1.092548f * y * z;
  Y[6] = 0.315392f * (3.0f * z * z - 1.0f);
  Y[7] = 1.092548f * x * z;
  Y[8] = 0.546274f * (x * x - y * y);
}

struct NeeResult {
  f3 L;        // radiance / joint pdf (light.cpp:185, 287-288)
  f3 dir;      // unnormalised shadow direction
  f3 origin;
  float tfar;
  int light;
  uint32_t tri;   // an area light's sampled triangle (for emitter_occludes)
};

// Point on an area-light triangle at folded barycentrics (u1, u2)
// (light.cpp:34-68): position v0 + u1 e1 + u2 e2 (fused; the reference
// blends w v0 + u1 v1 + u2 v2, the same point up to rounding), area pdf
// and emission texture coordinates.
__device__ __forceinline__ void area_point(const AreaTri& at, float u1, float u2, f3& target,
                                           float& ap, float& tu, float& tv) {
  target = {fmaf(u2, at.e2[0], fmaf(u1, at.e1[0], at.v0[0])),
            fmaf(u2, at.e2[1], fmaf(u1, at.e1[1], at.v0[1])),
            fmaf(u2, at.e2[2], fmaf(u1, at.e1[2], at.v0[2]))};
  ap = at.pdf;
  if (at.has_uv != 0.f) {
    float w = 1.0f - u1 - u2;
    tu = w * at.uv[0] + u1 * at.uv[2] + u2 * at.uv[4];
    tv = w * at.uv[1] + u1 * at.uv[3] + u2 * at.uv[5];
  }
}

// Fused forms of the light loop: x.x^2 + (x.y^2 + x.z^2) and a.b as
// x*x + (y*y + z*z) with the inner terms fused.
__device__ __forceinline__ float sqnorm_fma(f3 a) {
  return fmaf(a.x, a.x, fmaf(a.y, a.y, a.z * a.z));
}
__device__ __forceinline__ float dot_fma(f3 a, f3 b) {
  return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z));
}

// One NEE sample over all lights (light.cpp:82-189 at a path vertex when
// VERTEX, light.cpp:191-293 at a texel otherwise). Unshadowed radiance and a
// pending shadow ray per light (light.h:34-203; area lights sampled as in
// light.cpp:16-78), weight = maxCoeff, selection by a streaming reservoir:
// for each light draw u, W += w, keep light i when u*W < w -- the same
// w_i / sum(w) selection law as std::discrete_distribution. Returns kNeeNone
// when nothing is selected, sum(w) < 1e-6 or joint pdf < 1e-3, else
// kNeeShadow with the pending shadow ray in out (the caller may answer it
// with emitter_occludes instead of tracing it).
// Selection weights use the radiance without the light / albedo scales of
// shb_set_params, so the selection law does not depend on the fitted
// parameters (with every scale 1 this is the reference's law exactly); the
// returned radiance carries the scales. UNIT: every scale is 1, so the
// scaled and unscaled values are equal bit for bit (x * 1 == x, and brdf ==
// brdf_w) and the scale products are skipped.
// Light l at P: unshadowed radiance with the light's scale (rad) and
// without (radw, the selection weight's radiance), the pending shadow ray
// (direction, tfar) and the area pdf; an area light first draws its triangle
// (when it has several) and its barycentric pair from rng. VERTEX:
// BRDF x cosine applied (light.h *Radiance<Brdf>); else the texel form.
// The random numbers light L takes (an area light: its triangle when it has
// several, then its barycentric pair from one draw; other lights: none).
struct LightDraws {
  uint32_t tri;
  float u1, u2;
};
// A simple area light's draws and pick uniform (one state update).
__device__ __forceinline__ LightDraws simple_draws(Rng& rng, float& u) {
  LightDraws d{0u, 0.f, 0.f};
  rng.uniform_pair16_and_pick(d.u1, d.u2, u);
  return d;
}
// light_draws followed by the light's pick uniform of the one-sample NEE:
// an area light with triangles takes the uniform from its pair's step.
__device__ __forceinline__ LightDraws light_draws_pick(const LightDev& L, Rng& rng, float& u) {
  LightDraws d{0u, 0.f, 0.f};
  if (L.type == kLightArea && L.num_tris != 0) {
    d.tri = L.num_tris > 1 ? rng.index(L.num_tris) : 0u;
    rng.uniform_pair16_and_pick(d.u1, d.u2, u);
  } else {
    u = rng.uniform();
  }
  return d;
}
__device__ __forceinline__ LightDraws light_draws(const LightDev& L, Rng& rng) {
  LightDraws d{0u, 0.f, 0.f};
  if (L.type == kLightArea && L.num_tris != 0) {
    // single-triangle emitters draw no triangle index
    d.tri = L.num_tris > 1 ? rng.index(L.num_tris) : 0u;
    rng.uniform_pair16(d.u1, d.u2);
  }
  return d;
}

template <bool VERTEX, bool UNIT>
__device__ __forceinline__ void light_geom(const SceneView& sc, int l, const LightDev& L,
                                           const LightDraws& dr, f3 P, f3 N, f3 brdf,
                                           f3 brdf_w, bool front, f3& rad_out, f3& radw_out,
                                           f3& sdir_out, float& tfar_out, float& ap_out);

template <bool VERTEX, bool UNIT>
__device__ __forceinline__ void light_eval(const SceneView& sc, int l, f3 P, f3 N, f3 brdf,
                                           f3 brdf_w, bool front, Rng& rng, f3& rad_out,
                                           f3& radw_out, f3& sdir_out, float& tfar_out,
                                           float& ap_out) {
  // where every lane of the wave shares the job, the light record comes
  // through the scalar cache
  const LightDev L = const_load(sc.lights, (uint32_t)l);
  const LightDraws dr = light_draws(L, rng);
  light_geom<VERTEX, UNIT>(sc, l, L, dr, P, N, brdf, brdf_w, front, rad_out, radw_out, sdir_out,
                           tfar_out, ap_out);
}

// Light l's values from its draws (light_eval without the random numbers).
template <bool VERTEX, bool UNIT>
__device__ __forceinline__ void light_geom(const SceneView& sc, int l, const LightDev& L,
                                           const LightDraws& dr, f3 P, f3 N, f3 brdf,
                                           f3 brdf_w, bool front, f3& rad_out, f3& radw_out,
                                           f3& sdir_out, float& tfar_out, float& ap_out) {
  f3 rad = {0.f, 0.f, 0.f}, sdir;
  float tfar, ap = 1.0f, cos_n = 0.f;
  const int type = L.type;
  if (type == kLightDirectional) {
    f3 ld = {L.dir[0], L.dir[1], L.dir[2]};
    sdir = -ld;
    tfar = 1.0e10f;
    cos_n = dot(N, -ld);
    if (!(cos_n < 0.f)) rad = f3{L.ic[0], L.ic[1], L.ic[2]};
  } else {
    f3 target;
    f3 emit = {0.f, 0.f, 0.f};
    bool area = type == kLightArea;
    if (area) {
      if (L.num_tris == 0) {
        ap = 0.f;
        target = f3{0.f, 0.f, 0.f};
      } else {
        const uint32_t tri = dr.tri;
        float u1 = dr.u1;
        float u2 = dr.u2;
        if (u1 + u2 > 1.0f) {
          u1 = 1.0f - u1;
          u2 = 1.0f - u2;
        }
        // single-triangle emitters (tri is 0) read one record per job
        float tu = 0.f, tv = 0.f;
        if (L.num_tris == 1)
          area_point(const_load(sc.area_tris, L.first_tri), u1, u2, target, ap, tu, tv);
        else
          area_point(sc.area_tris[L.first_tri + tri], u1, u2, target, ap, tu, tv);
        if (L.emit_const)
          emit = f3{L.emit[0], L.emit[1], L.emit[2]};
        else
          emit = get_emission(sc, L.material, tu, tv);
      }
    } else {
      target = f3{L.pos[0], L.pos[1], L.pos[2]};
    }
    f3 Lv = target - P;
    float dist_sq = sqnorm_fma(Lv);
    float dist, inv;
    inv_dist(dist_sq, dist, inv);
    sdir = target - P;
    tfar = dist - 0.001f;
    // one division per light (the reference divides per component:
    // light.h:82-93, 127-143, 176-188) and fused sums; the oracle's
    // GPU-stream mode follows this arithmetic
    const float inv2 = inv * inv;
    Lv = Lv * inv;
    cos_n = dot_fma(N, Lv);
    if (type == kLightSpot) {
      f3 ld = {L.dir[0], L.dir[1], L.dir[2]};
      float cos_l = dot(ld, -Lv);
      float falloff = (cos_l - L.cos_outer) / (L.cos_inner - L.cos_outer);
      falloff = fminf(fmaxf(falloff, 0.0f), 1.0f);
      if (!(cos_n < 0.f)) rad = f3{L.ic[0], L.ic[1], L.ic[2]} * inv2 * falloff;
    } else if (type == kLightPoint) {
      if (!(cos_n < 0.f)) rad = f3{L.ic[0], L.ic[1], L.ic[2]} * inv2;
    } else {
      if (!(cos_n < 0.f)) rad = emit * inv2;
    }
  }
  // BRDF * cosine at a path vertex: scaled albedo for the contribution,
  // unscaled for the selection weight
  f3 radw = rad;
  if (VERTEX && !(cos_n < 0.f)) {
    bool ok = front && cos_n > 0.0f;
    f3 bv = ok ? brdf : f3{0.f, 0.f, 0.f};
    if (UNIT) {
      rad = cmul(rad, cos_n * bv);
      radw = rad;
    } else {
      f3 bw = ok ? brdf_w : f3{0.f, 0.f, 0.f};
      radw = cmul(rad, cos_n * bw);
      rad = cmul(rad, cos_n * bv);
    }
  }
  if (!UNIT) {
    const uint32_t l3 = 3u * (uint32_t)l;
    rad = cmul(rad, f3{const_load(sc.light_scale, l3), const_load(sc.light_scale, l3 + 1),
                       const_load(sc.light_scale, l3 + 2)});
  }
  rad_out = rad;
  radw_out = radw;
  sdir_out = sdir;
  tfar_out = tfar;
  ap_out = ap;
}

// Light L's selection weight maxc(radw) from its draws, as light_geom
// computes it (the same operations in the same order) but without what only
// the picked light needs: the shadow ray, the area pdf and the scaled
// radiance. The per-channel products are reduced with two identities that
// hold bit for bit because x -> fl(s x) never decreases for s >= 0: the
// texel weight max_c fl(k_c s) is fl(max_c k_c s), and at a path vertex a
// grey light's max_c fl(R fl(c b_c)) is fl(R fl(c max_c b_c)). bw is the
// BRDF of the weight (brdf_w, or brdf with unit scales) and maxbw its
// largest channel.
template <bool VERTEX>
__device__ __forceinline__ float light_weight(const SceneView& sc, const LightDev& L,
                                              const LightDraws& dr, f3 P, f3 N, f3 bw,
                                              float maxbw, bool front) {
  const int type = L.type;
  float cos_n;
  f3 k3 = {0.f, 0.f, 0.f};  // the radiance channels (non-grey vertex weight)
  float kmax;               // their largest (grey lights and texel weights)
  if (type == kLightDirectional) {
    const f3 ld = {L.dir[0], L.dir[1], L.dir[2]};
    cos_n = dot(N, -ld);
    if (cos_n < 0.f) return 0.f;
    kmax = L.wmax;
    if (VERTEX && !L.grey) k3 = f3{L.ic[0], L.ic[1], L.ic[2]};
  } else {
    f3 target;
    f3 emit = {0.f, 0.f, 0.f};
    float emax = L.wmax;
    if (type == kLightArea) {
      if (L.num_tris == 0) {
        target = f3{0.f, 0.f, 0.f};
      } else {
        float u1 = dr.u1, u2 = dr.u2;
        if (u1 + u2 > 1.0f) {
          u1 = 1.0f - u1;
          u2 = 1.0f - u2;
        }
        float ap, tu = 0.f, tv = 0.f;
        if (L.num_tris == 1)
          area_point(const_load(sc.area_tris, L.first_tri), u1, u2, target, ap, tu, tv);
        else
          area_point(sc.area_tris[L.first_tri + dr.tri], u1, u2, target, ap, tu, tv);
        if (L.emit_const) {
          emit = f3{L.emit[0], L.emit[1], L.emit[2]};
        } else {
          emit = get_emission(sc, L.material, tu, tv);
          emax = maxc(emit);
        }
      }
    } else {
      target = f3{L.pos[0], L.pos[1], L.pos[2]};
    }
    f3 Lv = target - P;
    const float dist_sq = sqnorm_fma(Lv);
    float dist, inv;
    inv_dist(dist_sq, dist, inv);
    const float inv2 = inv * inv;
    Lv = Lv * inv;
    cos_n = dot_fma(N, Lv);
    if (cos_n < 0.f) return 0.f;
    if (type == kLightSpot) {
      const f3 ld = {L.dir[0], L.dir[1], L.dir[2]};
      const float cos_l = dot(ld, -Lv);
      float falloff = (cos_l - L.cos_outer) / (L.cos_inner - L.cos_outer);
      falloff = fminf(fmaxf(falloff, 0.0f), 1.0f);
      kmax = emax * inv2 * falloff;
      if (VERTEX && !L.grey) k3 = f3{L.ic[0], L.ic[1], L.ic[2]} * inv2 * falloff;
    } else if (type == kLightPoint) {
      kmax = emax * inv2;
      if (VERTEX && !L.grey) k3 = f3{L.ic[0], L.ic[1], L.ic[2]} * inv2;
    } else {
      kmax = emax * inv2;
      if (VERTEX && !L.grey) k3 = emit * inv2;
    }
  }
  if (!VERTEX) return kmax;
  // BRDF x cosine, the BRDF zero unless the vertex faces the light
  const bool ok = front && cos_n > 0.0f;
  if (L.grey) return kmax * (cos_n * (ok ? maxbw : 0.f));
  return maxc(cmul(k3, cos_n * (ok ? bw : f3{0.f, 0.f, 0.f})));
}

// light_weight of a simple area light (one triangle, constant grey
// emitter: LightDev::simple_run > 0) from its draws: the same operations in
// the same order, without the branches on the light's kind.
template <bool VERTEX>
__device__ __forceinline__ float simple_area_weight(const AreaTri& at, float wmax, float u1,
                                                    float u2, f3 P, f3 N, float maxbw,
                                                    bool front) {
  if (u1 + u2 > 1.0f) {
    u1 = 1.0f - u1;
    u2 = 1.0f - u2;
  }
  const f3 target = {fmaf(u2, at.e2[0], fmaf(u1, at.e1[0], at.v0[0])),
                     fmaf(u2, at.e2[1], fmaf(u1, at.e1[1], at.v0[1])),
                     fmaf(u2, at.e2[2], fmaf(u1, at.e1[2], at.v0[2]))};
  f3 Lv = target - P;
  const float dist_sq = sqnorm_fma(Lv);
  float dist, inv;
  inv_dist(dist_sq, dist, inv);
  const float inv2 = inv * inv;
  Lv = Lv * inv;
  const float cos_n = dot_fma(N, Lv);
  const float kmax = wmax * inv2;
  float w = kmax;
  if (VERTEX) {
    const bool ok = front && cos_n > 0.0f;
    w = kmax * (cos_n * (ok ? maxbw : 0.f));
  }
  return cos_n < 0.f ? 0.f : w;
}

// Shadow query of an area light answered without traversal. The ray runs
// from P toward a point of one of the light's triangles, with t in units of
// |dir| (the point sits at t = 1), and reaches tfar = |dir| - 0.001
// (light.h:176-188): past the emitter whenever the light is farther than
// 1.001. The any-hit query (occlusion.cpp:88-107) is then occluded as soon as
// the watertight test accepts that triangle's own BVH record inside [tnear,
// tfar]: the traversal would report it or an earlier hit. emitter_occludes
// is the traversal's test (same ray setup, same triangle record, same
// arithmetic) and only accepts t in [0.5, 1.005] with tfar >= 1.01, well
// inside the range, where the traversal's widened box tests cannot cull the
// triangle; any other ray is traced.
// True when the shadow ray (origin o, tnear in o.w; direction d, tfar in
// d.w) toward triangle `tri` of light L is known to be occluded by that
// triangle (its copy of the BVH record, AreaTri::rec).
__device__ __forceinline__ bool emitter_occludes(const SceneView& sc, const LightDev& L,
                                                 uint32_t tri, float4 o, float4 d) {
  if (!sc.emitter_test || L.type != kLightArea || L.num_tris == 0 || !(d.w >= 1.01f) ||
      !(o.w <= 0.5f) || !ray_is_finite(o, d))
    return false;
  const AreaTri& at = sc.area_tris[L.first_tri + tri];
  if (at.leaf == kNoLeaf) return false;
  TravRay r;
  trav_setup(r, o, d);
  float rx, ry, rz;
  rot3(r.kz, r.ox, r.oy, r.oz, rx, ry, rz);
  float t, u, v;
  if (!tri_test(r.kz, rx, ry, rz, r.Sx, r.Sy, r.Sz, at.rec[0], at.rec[1], at.rec[2], t, u, v))
    return false;
  return t >= 0.5f && t <= 1.005f;
}

// nee_sample's outcomes: nothing picked (or pdf too small), a shadow ray, or
// (the caller's verdict) a pick whose shadow query emitter_occludes answered.
enum : uint32_t { kNeeNone = 0, kNeeShadow = 1, kNeeOccluded = 2 };

template <bool VERTEX, bool UNIT>
__device__ __forceinline__ uint32_t nee_sample(const SceneView& sc, f3 P, f3 N, f3 brdf,
                                               f3 brdf_w, bool front, Rng& rng,
                                               NeeResult& out) {
  const f3 bw = UNIT ? brdf : brdf_w;
  const float maxbw = maxc(bw);
  float W = 0.f, pw = 0.f;
  int pick = -1;
  LightDraws pd{0u, 0.f, 0.f};
  // the reservoir keeps the picked light's index, weight and draws; its
  // radiance, shadow ray and pdf are evaluated once after the loop
  auto take = [&](int l, const LightDraws& d, float w, float u) {
    W += w;
    if (w > 0.f && u * W < w) {
      pick = l;
      pw = w;
      pd = d;
    }
  };
  const int nl = sc.num_lights;
  // Lights in scene order. A run of simple area lights goes two per
  // iteration without branches on the kind: both lights' numbers are drawn
  // first, in the stream's order (the barycentric pair and pick uniform of
  // l from one step, then those of l + 1), so their weights are
  // independent work the scheduler can interleave; the reservoir then
  // takes l before l + 1. Other lights go one at a time.
  int l = 0;
  while (l < nl) {
    const int run = (int)const_load(sc.lights, (uint32_t)l).simple_run;
    if (run > 0) {
      const int end = l + run;
      for (; l + 1 < end; l += 2) {
        const LightDev La = const_load(sc.lights, (uint32_t)l);
        const LightDev Lb = const_load(sc.lights, (uint32_t)(l + 1));
        const AreaTri ta = const_load(sc.area_tris, La.first_tri);
        const AreaTri tb = const_load(sc.area_tris, Lb.first_tri);
        float ua, ub;
        const LightDraws da = simple_draws(rng, ua);
        const LightDraws db = simple_draws(rng, ub);
        const float wa = simple_area_weight<VERTEX>(ta, La.wmax, da.u1, da.u2, P, N, maxbw, front);
        const float wb = simple_area_weight<VERTEX>(tb, Lb.wmax, db.u1, db.u2, P, N, maxbw, front);
        take(l, da, wa, ua);
        take(l + 1, db, wb, ub);
      }
      if (l < end) {
        const LightDev L = const_load(sc.lights, (uint32_t)l);
        const AreaTri t = const_load(sc.area_tris, L.first_tri);
        float u;
        const LightDraws d = simple_draws(rng, u);
        take(l, d, simple_area_weight<VERTEX>(t, L.wmax, d.u1, d.u2, P, N, maxbw, front), u);
        ++l;
      }
    } else {
      const LightDev L = const_load(sc.lights, (uint32_t)l);
      float u;
      const LightDraws d = light_draws_pick(L, rng, u);
      take(l, d, light_weight<VERTEX>(sc, L, d, P, N, bw, maxbw, front), u);
      ++l;
    }
  }
  if (W < 1e-6f || pick < 0) return kNeeNone;
  f3 rad, radw, sdir;
  float tfar, ap;
  light_geom<VERTEX, UNIT>(sc, pick, sc.lights[pick], pd, P, N, brdf, brdf_w, front, rad, radw,
                           sdir, tfar, ap);
  float pdf = pw / W;
  float joint = pdf * ap;
  if (joint < 1e-3f) return kNeeNone;
  out.tri = pd.tri;
  out.L = rad / joint;
  out.dir = sdir;
  out.origin = P;
  out.tfar = tfar;
  out.light = pick;
  return kNeeShadow;
}

// n > 1 light samples (light.cpp:164-188 and 266-290): the reference's n
// picks with replacement from one set of per-light samples, as an inverse
// CDF. A first pass sums W (the area draws of every light), then n
// uniforms give targets u_j W, and a second pass replays the same area
// draws from a copy of the starting state and gives target j to the light
// whose interval of the running sum holds it -- light i with probability
// w_i / W, as discrete_distribution's. pick(j, result, joint) is called
// for every target whose joint pdf is at least 1e-3 (none when W < 1e-6),
// with result.L the light's radiance (not yet divided by the pdf), except
// targets whose light's shadow query ends on its own emitter
// (emitter_occludes): those are returned as a mask of occluded picks. rng
// ends after the targets (the material sample continues from there).
template <bool VERTEX, bool UNIT, class F>
__device__ __forceinline__ uint32_t nee_multi(const SceneView& sc, f3 P, f3 N, f3 brdf,
                                              f3 brdf_w, bool front, Rng& rng, int n,
                                              F&& pick) {
  const Rng start = rng;
  const int nl = sc.num_lights;
  const f3 bw = UNIT ? brdf : brdf_w;
  const float maxbw = maxc(bw);
  float W = 0.f;
  for (int l = 0; l < nl; ++l) {
    const LightDev L = const_load(sc.lights, (uint32_t)l);
    const LightDraws d = light_draws(L, rng);
    W += light_weight<VERTEX>(sc, L, d, P, N, bw, maxbw, front);
  }
  float target[kMaxLightSamples];
#pragma unroll
  for (int j = 0; j < kMaxLightSamples; ++j) target[j] = j < n ? rng.uniform() * W : 0.f;
  if (W < 1e-6f) return 0u;
  Rng replay = start;
  uint32_t open = (n >= 32 ? 0xffffffffu : (1u << n) - 1u);
  uint32_t occluded = 0u;
  float cum = 0.f;
  for (int l = 0; l < nl && open; ++l) {
    f3 rad, radw, sdir;
    float tfar, ap;
    const LightDev L = const_load(sc.lights, (uint32_t)l);
    const LightDraws dr = light_draws(L, replay);
    light_geom<VERTEX, UNIT>(sc, l, L, dr, P, N, brdf, brdf_w, front, rad, radw, sdir, tfar,
                             ap);
    const float w = maxc(radw);
    const float next = cum + w;
    // the targets in this light's interval (they share its sample and pdf)
    uint32_t here = 0u;
#pragma unroll
    for (int j = 0; j < kMaxLightSamples; ++j)
      if ((open & (1u << j)) && w > 0.f && target[j] < next) here |= 1u << j;
    open &= ~here;
    const float joint = (w / W) * ap;
    if (here && !(joint < 1e-3f)) {
      if (emitter_occludes(sc, L, dr.tri, make_float4(P.x, P.y, P.z, kRayEps),
                           make_float4(sdir.x, sdir.y, sdir.z, tfar))) {
        occluded |= here;
      } else {
        NeeResult r;
        r.L = rad;
        r.dir = sdir;
        r.origin = P;
        r.tfar = tfar;
        r.light = l;
        r.tri = dr.tri;
#pragma unroll
        for (int j = 0; j < kMaxLightSamples; ++j)
          if (here & (1u << j)) pick(j, r, joint);
      }
    }
    cum = next;
  }
  return occluded;
}

}  // namespace shb

#endif  // SHB_BAKE_DEVICE_H_
