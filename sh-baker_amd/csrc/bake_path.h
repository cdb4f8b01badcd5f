// Per-path bodies of the wavefront bake stages (generate, shade, project),
// shared by the HIP kernels and by CPU builds that check them. References
// are to DaviesX/sh-baker.
#ifndef SHB_BAKE_PATH_H_
#define SHB_BAKE_PATH_H_

#include "bake_device.h"
#include "kernels.h"

namespace shb {

// Primary ray and texel NEE of path slot p (baker.cpp:245-276,
// light.cpp:191-293). Returns true when a texel shadow ray was made (stored
// at shadow slot P + p).
__device__ __forceinline__ bool generate_path(const BakeArgs& a, uint32_t p) {
  const uint32_t T = a.texels_in_batch;
  uint32_t tl = p % T, smp = p / T;
  uint32_t texel = a.texel_list[a.texel_begin + tl];
  const float4* sp = a.surface + 3 * (size_t)texel;
  float4 q0 = sp[0], q1 = sp[1], q2 = sp[2];
  f3 P = {q0.x, q0.y, q0.z};
  f3 N = {q0.w, q1.x, q1.y};
  f3 T3 = {q1.z, q1.w, q2.x};
  float Tw = q2.y;
  f3 origin = P + N * 0.001f;
  Rng rng = site_rng(a.seed, texel, smp, 0, 1, kSiteTexel, a.job);
  float u1 = rng.uniform(), u2 = rng.uniform();
  f3 dl = hemisphere(u1, u2);
  f3 bit = cross(N, T3) * Tw;
  f3 dw = T3 * dl.x + bit * dl.y + N * dl.z;
  a.ray[2 * (size_t)p] = make_float4(origin.x, origin.y, origin.z, kRayEps);
  a.ray[2 * (size_t)p + 1] = make_float4(dw.x, dw.y, dw.z, INFINITY);
  a.dir0[p] = make_float4(dw.x, dw.y, dw.z, 0.f);
  a.thr[p] = make_float4(1.f, 1.f, 1.f, as_float(1u));
  a.li[p] = make_float4(0.f, 0.f, 0.f, 0.f);
  NeeResult nr;
  f3 zero = {0.f, 0.f, 0.f};
  if (nee_sample<false>(a.scene, P, N, zero, false, rng, nr)) {
    a.nee0[p] = make_float4(nr.L.x, nr.L.y, nr.L.z, 0.f);
    a.nee0_dir[p] = make_float4(nr.dir.x, nr.dir.y, nr.dir.z, 0.f);
    size_t so = (size_t)a.P + p;
    a.sray[2 * so] = make_float4(P.x, P.y, P.z, kRayEps);
    a.sray[2 * so + 1] = make_float4(nr.dir.x, nr.dir.y, nr.dir.z, nr.tfar);
    return true;
  }
  a.nee0[p] = make_float4(0.f, 0.f, 0.f, 0.f);
  return false;
}

// Shading of path slot p's closest hit at depth `depth` (baker.cpp:94-146):
// sky, emission at depth 0, vertex NEE (shadow slot p), cosine-weighted
// continuation. Sets shadow / extend when a ray was made.
__device__ __forceinline__ void shade_path(const BakeArgs& a, int depth, uint32_t p,
                                           bool& shadow, bool& extend) {
  shadow = extend = false;
  const uint32_t T = a.texels_in_batch;
  float4 h = a.hit[p];
  int id = (int)as_uint(h.y);
  if (id < 0) {
    if (depth == 0) a.li[p].w = 1.0f;  // sky seen directly (baker.cpp:83-91)
    return;
  }
  uint32_t tl = p % T, smp = p / T;
  uint32_t texel = a.texel_list[a.texel_begin + tl];
  float4 ro = a.ray[2 * (size_t)p], rd = a.ray[2 * (size_t)p + 1];
  f3 dir = f3_of(rd);
  f3 pos = f3_of(ro) + dir * h.x;
  int4 info = a.scene.tri_info[id];
  const int mat = info.x;
  const float4* sr = a.scene.shade + 4 * (size_t)id;
  float4 s0 = sr[0], s1 = sr[1], s2 = sr[2], s3 = sr[3];
  uint32_t fl = as_uint(s3.w);
  float u = h.z, v = h.w, w = 1.0f - u - v;
  f3 nrm;
  if (fl & 1u)
    nrm = normalized(w * f3_of(s0) + u * f3_of(s1) + v * f3_of(s2));
  else
    nrm = f3_of(s0);
  float tu = 0.f, tv = 0.f;
  if (fl & 2u) {
    tu = w * s0.w + u * s2.w + v * s3.y;
    tv = w * s1.w + u * s3.x + v * s3.z;
  }
  float4 th4 = a.thr[p];
  f3 thr = f3_of(th4);
  uint32_t branch = as_uint(th4.w);
  f3 add = {0.f, 0.f, 0.f};
  if (depth == 0 && a.scene.mats[mat].emission > 0.0f) {
    f3 em = get_emission(a.scene, mat, tu, tv);
    if (info.y >= 0) {
      const float* ls = a.scene.light_scale + 3 * info.y;
      em = cmul(em, f3{ls[0], ls[1], ls[2]});
    }
    add = cmul(thr, em);
  }
  f3 hit_pos = pos + nrm * 0.001f;
  f3 wo = -dir;
  bool front = dot(nrm, wo) > 0.0f;
  f3 brdf = lambert(a.scene, mat, tu, tv);
  Rng rng = site_rng(a.seed, texel, smp, (uint32_t)depth, branch, kSiteVertex, a.job);
  NeeResult nr;
  if (nee_sample<true>(a.scene, hit_pos, nrm, brdf, front, rng, nr)) {
    f3 c = cmul(thr, nr.L);
    a.scontrib[p] = make_float4(c.x, c.y, c.z, 0.f);
    a.sray[2 * (size_t)p] = make_float4(hit_pos.x, hit_pos.y, hit_pos.z, kRayEps);
    a.sray[2 * (size_t)p + 1] = make_float4(nr.dir.x, nr.dir.y, nr.dir.z, nr.tfar);
    shadow = true;
  }
  if (add.x != 0.f || add.y != 0.f || add.z != 0.f) {
    float4 l = a.li[p];
    l.x += add.x;
    l.y += add.y;
    l.z += add.z;
    a.li[p] = l;
  }
  float r1 = rng.uniform(), r2 = rng.uniform();
  f3 sd;
  float pdf;
  sample_lambert(nrm, r1, r2, sd, pdf);
  if (!(pdf < 1e-3f) && depth < (int)a.bounces) {
    // EvalMaterial is zero unless both directions are above the normal and
    // the albedo is non-zero; a zero BRDF ends the path (every deeper term
    // is zero) instead of tracing it.
    float cosine = dot(nrm, sd);
    if (front && cosine > 0.0f && !is_zero(brdf)) {
      f3 tn = cmul(thr, brdf) * (cosine / pdf);
      a.thr[p] = make_float4(tn.x, tn.y, tn.z, as_float(branch));
      a.ray[2 * (size_t)p] = make_float4(hit_pos.x, hit_pos.y, hit_pos.z, kRayEps);
      a.ray[2 * (size_t)p + 1] = make_float4(sd.x, sd.y, sd.z, INFINITY);
      extend = true;
    }
  }
}

// Folds texel tl's samples into SH9 (baker.cpp:284-294) with Neumaier-
// compensated FP32 sums, in the reference order per sample: texel NEE term
// along its unnormalised direction, then 2*pi*Li along the primary ray.
__device__ __forceinline__ void project_texel(const BakeArgs& a, uint32_t tl) {
  const uint32_t T = a.texels_in_batch;
  uint32_t texel = a.texel_list[a.texel_begin + tl];
  float sum[27], comp[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) sum[k] = comp[k] = 0.f;
  float vis = 0.f;
  const float two_pi = (float)(2.0 * kPi);  // baker.cpp:221
  auto addk = [&](int k, float x) {
    float t = sum[k] + x;
    if (fabsf(sum[k]) >= fabsf(x))
      comp[k] += (sum[k] - t) + x;
    else
      comp[k] += (x - t) + sum[k];
    sum[k] = t;
  };
  for (uint32_t smp = 0; smp < a.spp; ++smp) {
    uint32_t p = smp * T + tl;
    float4 n0 = a.nee0[p];
    float Y[9];
    if (n0.w != 0.f) {
      float4 nd = a.nee0_dir[p];
      sh_basis(f3_of(nd), Y);
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        addk(3 * i + 0, n0.x * Y[i]);
        addk(3 * i + 1, n0.y * Y[i]);
        addk(3 * i + 2, n0.z * Y[i]);
      }
    }
    float4 l = a.li[p];
    vis += l.w;
    f3 L = f3_of(l) * two_pi;
    if (L.x != 0.f || L.y != 0.f || L.z != 0.f) {
      sh_basis(f3_of(a.dir0[p]), Y);
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        addk(3 * i + 0, L.x * Y[i]);
        addk(3 * i + 1, L.y * Y[i]);
        addk(3 * i + 2, L.z * Y[i]);
      }
    }
  }
  const float inv = 1.0f / (float)a.spp;
  float* o = a.sh_out + 27 * (size_t)texel;
#pragma unroll
  for (int k = 0; k < 27; ++k) o[k] = (sum[k] + comp[k]) * inv;
  a.vis_out[texel] = vis * inv;
}

}  // namespace shb

#endif  // SHB_BAKE_PATH_H_
