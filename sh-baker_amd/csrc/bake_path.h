// Per-path bodies of the wavefront bake stages (generate, shade, project),
// shared by the HIP kernels and by CPU builds that check them. References
// are to DaviesX/sh-baker.
//
// A batch may hold several jobs (lightmaps); a path slot finds its job,
// texel and sample through the batch's BatchMap. When every active lane of
// a wave works on the same job, the job record is read as a wave-uniform
// value (scalar loads); otherwise per lane.
#ifndef SHB_BAKE_PATH_H_
#define SHB_BAKE_PATH_H_

#include "bake_device.h"
#include "kernels.h"

namespace shb {

// Where path slot p of the batch lives.
struct SlotRef {
  uint32_t job;
  uint32_t rank;  // texel rank in the job's texel list
  uint32_t smp;
};

__device__ __forceinline__ SlotRef slot_ref(const BakeArgs& a, uint32_t p) {
  SlotRef r;
  if (a.nseg <= 1) {
    r.job = a.job0;
    r.rank = a.texel_begin0 + p % a.T0;
    r.smp = p / a.T0;
    return r;
  }
  const BatchMap& m = *a.map;
  const uint32_t s = find_segment(m.slot_base, a.nseg, p);
  const uint32_t Ts = m.tex_base[s + 1] - m.tex_base[s];
  const uint32_t rel = p - m.slot_base[s];
  r.job = m.job[s];
  r.rank = m.texel_begin[s] + rel % Ts;
  r.smp = rel / Ts;
  return r;
}

// The sample (root) slot of path slot p: itself, or for a branch slot the
// root recorded when it was spawned.
__device__ __forceinline__ uint32_t root_slot(const BakeArgs& a, uint32_t p) {
  return p < a.P ? p : a.link[p - a.P].x;
}

// Calls f(job record) with the record read wave-uniformly when all active
// lanes share `job` (the common case: segments are large and contiguous).
// ONE: the batch is one segment, so every slot is job0's and f is compiled
// once, with the job's tables and light records in scalar registers.
template <bool ONE, class F>
__device__ __forceinline__ void with_job(const BakeArgs& a, uint32_t job, F&& f) {
#if defined(SHB_HOST_ONLY)
  f(a.jobs[job]);
#else
  if (ONE) {
    const JobDev J = uniform_load(a.jobs, a.job0);
    f(J);
    return;
  }
  const uint32_t j0 = uniform_index(job);
  if (__ballot(job != j0) == 0ull) {
    const JobDev J = uniform_load(a.jobs, j0);
    f(J);
  } else {
    const JobDev J = a.jobs[job];
    f(J);
  }
#endif
}

// Primary ray and texel NEE of path slot p (baker.cpp:245-276,
// light.cpp:191-293). Returns the mask of texel shadow rays made: bit j for
// the ray of pick j at shadow slot nls * Pc + j * P + p (nls = 1: bit 0,
// slot Pc + p); bit 16 + j when pick j's shadow query was answered occluded
// without a ray (emitter_occludes). UNIT: every light and albedo scale is 1. MULTI: the batch
// has a job with num_light_samples > 1 (per-pick arrays, visible flags).
template <bool UNIT, bool MULTI>
__device__ __forceinline__ uint32_t generate_path(const BakeArgs& a, const JobDev& J,
                                                  const SlotRef& sr, uint32_t p) {
  uint32_t texel = J.texel_list[sr.rank];
  const float4* sp = J.surface + 3 * (size_t)texel;
  float4 q0 = sp[0], q1 = sp[1], q2 = sp[2];
  f3 P = {q0.x, q0.y, q0.z};
  f3 N = {q0.w, q1.x, q1.y};
  f3 T3 = {q1.z, q1.w, q2.x};
  float Tw = q2.y;
  f3 origin = P + N * 0.001f;
  Rng rng = site_rng(J.seed, texel, sr.smp, 0, 1, kSiteTexel, J.stream);
  float u1 = rng.uniform(), u2 = rng.uniform();
  f3 dl = hemisphere(u1, u2);
  f3 bit = cross(N, T3) * Tw;
  f3 dw = T3 * dl.x + bit * dl.y + N * dl.z;
  a.ray[2 * (size_t)p] = make_float4(origin.x, origin.y, origin.z, kRayEps);
  a.ray[2 * (size_t)p + 1] = make_float4(dw.x, dw.y, dw.z, INFINITY);
  a.dir0[p] = make_float4(dw.x, dw.y, dw.z, 0.f);
  if (a.adj_grad) {
    // d SH[k][c] / d Li_c = 2 pi Y_k(dir0) / spp (baker.cpp:286-292)
    const float* g = a.adj_dsh + 27 * (size_t)texel;
    float Y[9];
    sh_basis(dw, Y);
    float gr = 0.f, gg = 0.f, gb = 0.f;
    for (int k = 0; k < 9; ++k) {
      gr += g[3 * k + 0] * Y[k];
      gg += g[3 * k + 1] * Y[k];
      gb += g[3 * k + 2] * Y[k];
    }
    const float f = (float)(2.0 * kPi) * (1.0f / (float)J.spp);
    a.adj_gvec[p] = make_float4(gr * f, gg * f, gb * f, 0.f);
  }
  a.thr[p] = make_float4(1.f, 1.f, 1.f, as_float(1u));
  a.li[p] = make_float4(0.f, 0.f, 0.f, 0.f);
  NeeResult nr;
  f3 zero = {0.f, 0.f, 0.f};
  if (MULTI) {
    // one entry per pick, invisible until the shadow trace flags it
    for (uint32_t j = 0; j < a.nls; ++j)
      a.nee0[(size_t)j * a.P + p] = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t made = 0;
    auto put = [&](int j, const f3& Li, const NeeResult& r) {
      const size_t e = (size_t)j * a.P + p;
      a.nee0[e] = make_float4(Li.x, Li.y, Li.z, 0.f);
      a.nee0_dir[e] = make_float4(r.dir.x, r.dir.y, r.dir.z, as_float((uint32_t)r.light));
      const size_t so = (size_t)a.nls * a.Pc + e;
      a.sray[2 * so] = make_float4(P.x, P.y, P.z, kRayEps);
      a.sray[2 * so + 1] = make_float4(r.dir.x, r.dir.y, r.dir.z, r.tfar);
      made |= 1u << j;
    };
    const int n = (int)J.nls;
    uint32_t occluded = 0u;
    if (n > 1) {
      // Li = L / (joint n) per visible pick (light.cpp:287-288)
      occluded = nee_multi<false, UNIT>(J.scene, P, N, zero, zero, false, rng, n,
                                        [&](int j, const NeeResult& r, float joint) {
                                          put(j, r.L / (joint * (float)n), r);
                                        });
    } else if (nee_sample<false, UNIT>(J.scene, P, N, zero, zero, false, rng, nr) == kNeeShadow) {
      if (emitter_occludes(J.scene, J.scene.lights[nr.light], nr.tri,
                           make_float4(P.x, P.y, P.z, kRayEps),
                           make_float4(nr.dir.x, nr.dir.y, nr.dir.z, nr.tfar)))
        occluded = 1u;
      else
        put(0, nr.L, nr);
    }
    return made | occluded << 16;
  }
  uint32_t ns = nee_sample<false, UNIT>(J.scene, P, N, zero, zero, false, rng, nr);
  if (ns == kNeeShadow &&
      emitter_occludes(J.scene, J.scene.lights[nr.light], nr.tri,
                       make_float4(P.x, P.y, P.z, kRayEps),
                       make_float4(nr.dir.x, nr.dir.y, nr.dir.z, nr.tfar)))
    ns = kNeeOccluded;
  if (ns == kNeeShadow) {
    a.nee0[p] = make_float4(nr.L.x, nr.L.y, nr.L.z, 0.f);
    a.nee0_dir[p] = make_float4(nr.dir.x, nr.dir.y, nr.dir.z, as_float((uint32_t)nr.light));
    size_t so = (size_t)a.Pc + p;
    a.sray[2 * so] = make_float4(P.x, P.y, P.z, kRayEps);
    a.sray[2 * so + 1] = make_float4(nr.dir.x, nr.dir.y, nr.dir.z, nr.tfar);
    return 1u;
  }
  a.nee0[p] = make_float4(0.f, 0.f, 0.f, 0.f);
  return ns == kNeeOccluded ? 1u << 16 : 0u;
}

// Appends a wave's spawning lanes to the batch's branch slots (one atomic
// per wave); kNoSlot when the spare slots have run out.
__device__ __forceinline__ uint32_t claim_branch_slot(const BakeArgs& a) {
#if defined(SHB_HOST_ONLY)
  const uint32_t k = (*a.spawn_count)++;
#else
  const uint64_t m = __ballot(true);
  const int leader = __ffsll((unsigned long long)m) - 1;
  const uint32_t lane = __lane_id();
  uint32_t base = 0;
  if ((int)lane == leader) base = atomicAdd(a.spawn_count, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  const uint32_t k = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
#endif
  return k < a.S ? k : kNoSlot;
}

// Shading of path slot p's closest hit at depth `depth` (baker.cpp:94-146):
// sky, alpha split, emission at depth 0, vertex NEE (shadow slot p; pick j
// at j * Pc + p), cosine-weighted continuation. Sets `shadow` to the mask of
// vertex shadow rays made (bit j: pick j; bit 16 + j: pick j's query answered
// occluded without a ray; with one light sample the caller tests the stored
// ray with emitter_occludes instead), `extend` when a continuation ray
// was made and `child` to the branch slot of a transmitted ray (kNoSlot:
// none). `sr` is the root slot's texel and sample. MULTI: per-pick
// contributions with visible flags (the batch has a job with
// num_light_samples > 1); every entry of the slot is cleared first, so
// k_nee_gather reads zeros for picks not made at this depth.
// ADJ: records the vertex's material for the replay adjoint and adds the
// gradient of the depth-0 emission term to `acc`.
// SPLIT: the batch has spare slots (S > 0). Without them no material has a
// texel with alpha < 255 (the host sets S from the texel pool), every alpha
// is 1 and the split code is compiled out. UNIT: every light and albedo
// scale is 1.
template <bool ADJ, bool SPLIT, bool UNIT, bool MULTI>
__device__ __forceinline__ void shade_path(const BakeArgs& a, const JobDev& J, const SlotRef& sr,
                                           int depth, uint32_t p, uint32_t& shadow, bool& extend,
                                           uint32_t& child, GradAcc* acc) {
  shadow = 0u;
  extend = false;
  child = kNoSlot;
  if (MULTI)
    for (uint32_t j = 0; j < a.nls; ++j)
      a.scontrib[(size_t)j * a.Pc + p] = make_float4(0.f, 0.f, 0.f, 0.f);
  const SceneView& sc = J.scene;
  float4 h = a.hit[p];
  int id = (int)as_uint(h.y);
  if (id < 0) {
    if (depth == 0) a.li[p].w = 1.0f;  // sky seen directly (baker.cpp:83-91)
    return;
  }
  uint32_t texel = J.texel_list[sr.rank];
  float4 ro = a.ray[2 * (size_t)p], rd = a.ray[2 * (size_t)p + 1];
  f3 dir = f3_of(rd);
  f3 pos = f3_of(ro) + dir * h.x;
  int4 info = sc.tri_info[id];
  const int mat = info.x;
  const float4* srec = sc.shade + 4 * (size_t)id;
  float4 s0 = srec[0], s1 = srec[1], s2 = srec[2], s3 = srec[3];
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
  uint32_t next_branch = branch;
  // alpha < 1 (baker.cpp:97-109): the path splits into the transmitted ray
  // from pos + 0.001 dir, weight 1 - alpha, and the reflected part
  // (emission, NEE, continuation), weight alpha; alpha 0 keeps only the
  // transmitted ray. Branch ids: transmitted 2b + 1, reflected continuation
  // 2b. The transmitted ray goes on in this slot when alpha is 0, else in a
  // spare slot of the batch; no ray is made past the last bounce
  // (baker.cpp:73).
  const float alpha = SPLIT ? get_alpha(sc, mat, tu, tv) : 1.0f;
  if (SPLIT && alpha < 1.0f) {
    const f3 o = pos + dir * 0.001f;
    const uint32_t tb = 2u * branch + 1u;
    if (alpha == 0.0f) {
      if (ADJ) a.adj_hist[(size_t)depth * a.Pc + p] = kHistTransmit;
      if (depth < (int)J.bounces) {
        a.thr[p] = make_float4(thr.x, thr.y, thr.z, as_float(tb));
        a.ray[2 * (size_t)p] = make_float4(o.x, o.y, o.z, kRayEps);
        a.ray[2 * (size_t)p + 1] = make_float4(dir.x, dir.y, dir.z, INFINITY);
        extend = true;
      }
      return;
    }
    if (depth < (int)J.bounces) {
      const uint32_t k = claim_branch_slot(a);
      if (k != kNoSlot) {
        const uint32_t c = a.P + k;
        const f3 tt = thr * (1.0f - alpha);
        a.link[k] = make_uint4(root_slot(a, p), p, (uint32_t)depth, 0u);
        a.thr[c] = make_float4(tt.x, tt.y, tt.z, as_float(tb));
        a.li[c] = make_float4(0.f, 0.f, 0.f, 0.f);
        a.ray[2 * (size_t)c] = make_float4(o.x, o.y, o.z, kRayEps);
        a.ray[2 * (size_t)c + 1] = make_float4(dir.x, dir.y, dir.z, INFINITY);
        if (ADJ) {
          // the branch carries the parent's vertices before this one
          a.adj_gvec[c] = a.adj_gvec[p];
          for (int k2 = 0; k2 < depth; ++k2)
            a.adj_hist[(size_t)k2 * a.Pc + c] = a.adj_hist[(size_t)k2 * a.Pc + p];
          a.adj_hist[(size_t)depth * a.Pc + c] = kHistTransmit;
        }
        child = c;
      }
    }
    thr = thr * alpha;
    next_branch = 2u * branch;
  }
  f3 add = {0.f, 0.f, 0.f};
  if (ADJ) a.adj_hist[(size_t)depth * a.Pc + p] = (uint16_t)mat;
  if (depth == 0 && sc.mats[mat].emission > 0.0f) {
    f3 em = get_emission(sc, mat, tu, tv);
    if (info.y >= 0) {
      const float* ls = sc.light_scale + 3 * info.y;
      if (ADJ) {
        // d/d scale of the emitter: unscaled emission times throughput (1 at
        // depth 0, alpha after a split) times the projection weight
        const float4 g = a.adj_gvec[p];
        acc->add(3u * info.y + 0, em.x * thr.x * g.x);
        acc->add(3u * info.y + 1, em.y * thr.y * g.y);
        acc->add(3u * info.y + 2, em.z * thr.z * g.z);
      }
      em = cmul(em, f3{ls[0], ls[1], ls[2]});
    }
    add = cmul(thr, em);
  }
  f3 hit_pos = pos + nrm * 0.001f;
  f3 wo = -dir;
  bool front = dot(nrm, wo) > 0.0f;
  f3 brdf_w;
  f3 brdf = lambert(sc, mat, tu, tv, &brdf_w);
  Rng rng = site_rng(J.seed, texel, sr.smp, (uint32_t)depth, branch, kSiteVertex, J.stream);
  NeeResult nr;
  if (MULTI) {
    // pick j: contribution thr x L / joint, visible flag in w (0 until the
    // shadow trace); k_nee_gather adds the visible ones / n
    auto put = [&](int j, const f3& Lj, const NeeResult& r) {
      const f3 c = cmul(thr, Lj);
      const size_t e = (size_t)j * a.Pc + p;
      a.scontrib[e] = make_float4(c.x, c.y, c.z, 0.f);
      a.sray[2 * e] = make_float4(hit_pos.x, hit_pos.y, hit_pos.z, kRayEps);
      a.sray[2 * e + 1] = make_float4(r.dir.x, r.dir.y, r.dir.z, r.tfar);
      shadow |= 1u << j;
    };
    const int n = (int)J.nls;
    if (n > 1) {
      const uint32_t occluded =
          nee_multi<true, UNIT>(sc, hit_pos, nrm, brdf, brdf_w, front, rng, n,
                                [&](int j, const NeeResult& r, float joint) {
                                  put(j, r.L / joint, r);
                                });
      shadow |= occluded << 16;
    } else if (nee_sample<true, UNIT>(sc, hit_pos, nrm, brdf, brdf_w, front, rng, nr) ==
               kNeeShadow) {
      if (emitter_occludes(sc, sc.lights[nr.light], nr.tri,
                           make_float4(hit_pos.x, hit_pos.y, hit_pos.z, kRayEps),
                           make_float4(nr.dir.x, nr.dir.y, nr.dir.z, nr.tfar)))
        shadow |= 1u << 16;
      else
        put(0, nr.L, nr);
    }
  } else if (nee_sample<true, UNIT>(sc, hit_pos, nrm, brdf, brdf_w, front, rng, nr) ==
             kNeeShadow) {
    // the shadow ray is stored; the caller may still answer it with
    // emitter_occludes once the path's other values are dead
    f3 c = cmul(thr, nr.L);
    a.scontrib[p] = make_float4(c.x, c.y, c.z, as_float((uint32_t)nr.light));
    a.sray[2 * (size_t)p] = make_float4(hit_pos.x, hit_pos.y, hit_pos.z, kRayEps);
    a.sray[2 * (size_t)p + 1] = make_float4(nr.dir.x, nr.dir.y, nr.dir.z, nr.tfar);
    shadow = 1u;
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
  if (!(pdf < 1e-3f) && depth < (int)J.bounces) {
    // EvalMaterial is zero unless both directions are above the normal and
    // the albedo is non-zero; a zero BRDF ends the path (every deeper term
    // is zero) instead of tracing it.
    float cosine = dot(nrm, sd);
    if (front && cosine > 0.0f && !is_zero(brdf)) {
      f3 tn = cmul(thr, brdf) * (cosine / pdf);
      a.thr[p] = make_float4(tn.x, tn.y, tn.z, as_float(next_branch));
      a.ray[2 * (size_t)p] = make_float4(hit_pos.x, hit_pos.y, hit_pos.z, kRayEps);
      a.ray[2 * (size_t)p + 1] = make_float4(sd.x, sd.y, sd.z, INFINITY);
      extend = true;
    }
  }
}

// Adjoint of a texel's NEE terms: for each visible sample, the light's
// gradient gains (unscaled radiance) x (1/spp) sum_k dL/dSH[k][c] Y_k(d).
__device__ __forceinline__ void project_texel_adjoint(const BakeArgs& a, const JobDev& J,
                                                      uint32_t rank, uint32_t slot0,
                                                      uint32_t stride, GradAcc* acc) {
  uint32_t texel = J.texel_list[rank];
  const float* g = a.adj_dsh + 27 * (size_t)texel;
  const float inv = 1.0f / (float)J.spp;
  for (uint32_t smp = 0; smp < J.spp; ++smp) {
    uint32_t p = slot0 + smp * stride;
    float4 n0 = a.nee0[p];
    if (n0.w == 0.f) continue;
    float4 nd = a.nee0_dir[p];
    const uint32_t l = as_uint(nd.w);
    float Y[9];
    sh_basis(f3_of(nd), Y);
    float gr = 0.f, gg = 0.f, gb = 0.f;
    for (int k = 0; k < 9; ++k) {
      gr += g[3 * k + 0] * Y[k];
      gg += g[3 * k + 1] * Y[k];
      gb += g[3 * k + 2] * Y[k];
    }
    const float* s = J.scene.light_scale + 3 * l;
    acc->add(3 * l + 0, n0.x / s[0] * gr * inv);
    acc->add(3 * l + 1, n0.y / s[1] * gg * inv);
    acc->add(3 * l + 2, n0.z / s[2] * gb * inv);
  }
}

// Adjoint of slot p's visible vertex NEE term at depth d: value
// (throughput x radiance, per channel) times the projection weight, divided
// by the light's scale for the light gradient and by each path vertex's
// albedo scale (vertices 0..d) for the material gradients.
__device__ __forceinline__ void adjoint_vertex(const BakeArgs& a, const JobDev& J, int depth,
                                               uint32_t p, GradAcc* acc) {
  const float4 c = a.scontrib[p];
  const float4 g = a.adj_gvec[p];
  const float G[3] = {c.x * g.x, c.y * g.y, c.z * g.z};
  const uint32_t l = as_uint(c.w);
  const float* s = J.scene.light_scale + 3 * l;
  for (int ch = 0; ch < 3; ++ch) acc->add(3 * l + ch, G[ch] / s[ch]);
  const uint32_t mbase = 3 * a.adj_nlights;
  for (int k = 0; k <= depth; ++k) {
    const uint32_t m = a.adj_hist[(size_t)k * a.Pc + p];
    if (m == kHistTransmit) continue;  // a transmitted vertex has no albedo factor
    const float* as = J.scene.albedo_scale + 3 * m;
    for (int ch = 0; ch < 3; ++ch) acc->add(mbase + 3 * m + ch, G[ch] / as[ch]);
  }
}

// Folds batch texel t's samples into SH9 (baker.cpp:284-294) with Neumaier-
// compensated FP32 sums, in the reference order per sample: texel NEE term
// along its unnormalised direction, then 2*pi*Li along the primary ray.
__device__ __forceinline__ void project_texel(const BakeArgs& a, const JobDev& J, uint32_t rank,
                                              uint32_t slot0, uint32_t stride) {
  uint32_t texel = J.texel_list[rank];
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
  for (uint32_t smp = 0; smp < J.spp; ++smp) {
    uint32_t p = slot0 + smp * stride;
    float Y[9];
    // the texel NEE picks in order (one unless num_light_samples > 1)
    for (uint32_t j = 0; j < a.nls; ++j) {
      const size_t e = (size_t)j * a.P + p;
      float4 n0 = a.nee0[e];
      if (n0.w != 0.f) {
        float4 nd = a.nee0_dir[e];
        sh_basis(f3_of(nd), Y);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
          addk(3 * i + 0, n0.x * Y[i]);
          addk(3 * i + 1, n0.y * Y[i]);
          addk(3 * i + 2, n0.z * Y[i]);
        }
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
  const float inv = 1.0f / (float)J.spp;
  float* o = J.sh_out + 27 * (size_t)texel;
#pragma unroll
  for (int k = 0; k < 27; ++k) o[k] = (sum[k] + comp[k]) * inv;
  J.vis_out[texel] = vis * inv;
}

}  // namespace shb

#endif  // SHB_BAKE_PATH_H_
