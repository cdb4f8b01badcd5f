// Wavefront path tracer for the SH9 lightmap bake on CDNA4.
//
// A batch holds T texels x spp samples as path slots p = sample * T + texel
// (sample-major, so the SH projection reads coalesced). Per bounce depth d:
//   k_trace<closest>  over the extension queue        (rtcIntersect1)
//   k_shade           hit -> emission, NEE pick, material sample, next ray
//   k_trace<any>      over the shadow queue, epilogue adds the NEE term
// k_generate makes the primary rays and the texel NEE shadow rays; k_project
// folds each texel's samples into 27 SH coefficients with compensated sums.
// Mirrors BakeSHLightMap / Trace / EvaluateLightSamples /
// AccumulateIncomingLightSamples (reference src/baker.cpp:70-304,
// src/light.cpp:82-293).
#include "../../include/shb.h"
#include "bake_device.h"
#include "kernels.h"
#include "traverse.h"

namespace shb {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Appends `pred` lanes to a queue with one atomic per wave. Every lane of the
// wave must call it. Returns the slot of a predicated lane.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
  uint64_t m = __ballot(pred);
  if (m == 0) return 0u;
  int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if ((int)lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  return base + (uint32_t)__popcll(m & ((1ull << lane_id()) - 1ull));
}

// ------------------------------------------------------------------ trace
struct HitEpilogue {
  float4* hits;
  __device__ void operator()(uint32_t slot, const TravState& s) const {
    float4 h;
    h.x = s.tbest;
    h.y = as_float(s.best == 0x7fffffff ? 0xffffffffu : (uint32_t)s.best);
    h.z = s.bu;
    h.w = s.bv;
    hits[slot] = h;
  }
};

struct OccludedEpilogue {
  uint32_t* occluded;
  __device__ void operator()(uint32_t slot, const TravState& s) const {
    occluded[slot] = s.best != 0x7fffffff ? 1u : 0u;
  }
};

// Shadow rays of the bake: slot < P is a vertex NEE ray (adds its
// contribution to the path's radiance), slot >= P a texel NEE ray (marks the
// texel term visible).
struct BakeShadowEpilogue {
  float4* li;
  const float4* contrib;
  float4* nee0;
  uint32_t P;
  __device__ void operator()(uint32_t slot, const TravState& s) const {
    if (s.best != 0x7fffffff) return;
    if (slot < P) {
      float4 c = contrib[slot];
      float4 l = li[slot];
      l.x += c.x;
      l.y += c.y;
      l.z += c.z;
      li[slot] = l;
    } else {
      nee0[slot - P].w = 1.0f;
    }
  }
};

// Persistent traversal: each wave keeps 64 rays in flight and refills lanes
// whose ray finished from the queue (one atomic per refill), so a wave does
// not idle on its slowest ray (Aila & Laine, HPG 2009, dynamic fetch).
template <bool ANY, bool COUNT, class Epi>
__global__ void __launch_bounds__(256) k_trace(TraceArgs a, Epi epi) {
  extern __shared__ uint2 lds_stack[];
  const uint32_t count = a.count_ptr ? *a.count_ptr : a.count_const;
  const int tid = threadIdx.x;
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  uint2* lds = lds_stack + tid;
  uint2* spill = a.spill ? a.spill + (size_t)gtid * a.spill_cap : nullptr;
  const int stack_cap = kStackLds + a.spill_cap;
  bool active = false;
  bool exhausted = false;
  TravRay r;
  TravState s;
  uint32_t slot = 0;
  uint64_t cnt_nodes = 0, cnt_tris = 0;
  for (;;) {
    uint64_t inact = __ballot(!active);
    int ninact = __popcll(inact);
    if (!exhausted && (ninact >= a.refill || ninact == 64)) {
      uint32_t base = 0;
      if (lane_id() == 0) base = atomicAdd(a.work, (uint32_t)ninact);
      base = __shfl(base, 0);
      if (!active) {
        uint32_t my = base + (uint32_t)__popcll(inact & ((1ull << lane_id()) - 1ull));
        if (my < count) {
          slot = a.queue ? a.queue[my] : my;
          float4 o = a.rays[2 * (size_t)slot];
          float4 d = a.rays[2 * (size_t)slot + 1];
          trav_setup(r, o, d);
          trav_init(s, r, d.w);
          if (!a.has_nodes) s.gbits = 0;  // empty scene: every ray misses
          active = true;
        }
      }
      if (base + (uint32_t)ninact >= count) exhausted = true;
    }
    if (__ballot(active) == 0) {
      if (exhausted) break;
      continue;
    }
    for (int it = 0; it < a.iters; ++it) {
      if (!active) break;
      bool overflow = false;
      bool done = trav_step<ANY, COUNT>(a.scene, r, s, lds, blockDim.x, spill,
                                        overflow, stack_cap);
      if (done) {
        if (overflow) {
          atomicOr(a.error, 1u);
          s.best = 0x7fffffff;
        }
        if (COUNT) {
          if (a.ray_counts) {
            a.ray_counts[2 * (size_t)slot] = s.nodes;
            a.ray_counts[2 * (size_t)slot + 1] = s.tris;
          }
          cnt_nodes += s.nodes;
          cnt_tris += s.tris;
        }
        epi(slot, s);
        active = false;
      }
    }
  }
  if (COUNT && a.totals) {
    atomicAdd(a.totals, (unsigned long long)cnt_nodes);
    atomicAdd(a.totals + 1, (unsigned long long)cnt_tris);
  }
}

// ------------------------------------------------------------- fixed rays
// Fills shb_hit from a closest-hit record (occlusion.cpp:39-104).
__global__ void k_hit_details(SceneView sc, const float4* rays, const float4* hits,
                              uint32_t n, shb_hit* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 h = hits[i];
  shb_hit o;
  int id = (int)as_uint(h.y);
  o.t = h.x;
  o.u = h.z;
  o.v = h.w;
  o.prim_id = id;
  o.geom_id = -1;
  o.material_id = -1;
  for (int k = 0; k < 3; ++k) o.position[k] = o.normal[k] = 0.f;
  o.uv[0] = o.uv[1] = 0.f;
  if (id >= 0) {
    float4 ro = rays[2 * (size_t)i], rd = rays[2 * (size_t)i + 1];
    f3 pos = f3_of(ro) + f3_of(rd) * h.x;
    int4 info = sc.tri_info[id];
    const float4* sr = sc.shade + 4 * (size_t)id;
    float4 s0 = sr[0], s1 = sr[1], s2 = sr[2], s3 = sr[3];
    uint32_t fl = as_uint(s3.w);
    float u = h.z, v = h.w, w = 1.0f - u - v;
    f3 n;
    if (fl & 1u)
      n = normalized(w * f3_of(s0) + u * f3_of(s1) + v * f3_of(s2));
    else
      n = f3_of(s0);
    float tu = 0.f, tv = 0.f;
    if (fl & 2u) {
      tu = w * s0.w + u * s2.w + v * s3.y;
      tv = w * s1.w + u * s3.x + v * s3.z;
    }
    o.geom_id = info.z;
    o.material_id = info.x;
    o.position[0] = pos.x;
    o.position[1] = pos.y;
    o.position[2] = pos.z;
    o.normal[0] = n.x;
    o.normal[1] = n.y;
    o.normal[2] = n.z;
    o.uv[0] = tu;
    o.uv[1] = tv;
  }
  out[i] = o;
}

// ------------------------------------------------------------ bake stages
// Primary ray + texel NEE for every path slot of the batch
// (baker.cpp:245-276; light.cpp:191-293).
__global__ void k_generate(BakeArgs a) {
  const uint32_t T = a.texels_in_batch;
  const uint32_t n = T * a.spp;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    uint32_t p = base + threadIdx.x;
    bool valid = p < n;
    bool shadow = false;
    if (valid) {
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
        shadow = true;
      } else {
        a.nee0[p] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
    uint32_t at = wave_append(a.shadow_count, shadow);
    if (shadow) a.shadow_queue[at] = a.P + p;
  }
}

// Hit shading at depth d (baker.cpp:94-146).
__global__ void k_shade(BakeArgs a, int depth, const uint32_t* in_queue,
                        const uint32_t* in_count, uint32_t* out_queue,
                        uint32_t* out_count) {
  const uint32_t T = a.texels_in_batch;
  const uint32_t n = in_queue ? *in_count : T * a.spp;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    uint32_t i = base + threadIdx.x;
    bool valid = i < n;
    bool shadow = false, extend = false;
    uint32_t p = 0;
    if (valid) {
      p = in_queue ? in_queue[i] : i;
      float4 h = a.hit[p];
      int id = (int)as_uint(h.y);
      if (id < 0) {
        if (depth == 0) a.li[p].w = 1.0f;  // sky seen directly (baker.cpp:83-91)
      } else {
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
        if (!(pdf < 1e-3f) && depth < a.bounces) {
          // EvalMaterial: zero unless both directions are above the normal;
          // a zero BRDF ends the path (its deeper terms are all zero).
          float cosine = dot(nrm, sd);
          if (front && cosine > 0.0f) {
            f3 tn = cmul(thr, brdf) * (cosine / pdf);
            a.thr[p] = make_float4(tn.x, tn.y, tn.z, as_float(branch));
            a.ray[2 * (size_t)p] = make_float4(hit_pos.x, hit_pos.y, hit_pos.z, kRayEps);
            a.ray[2 * (size_t)p + 1] = make_float4(sd.x, sd.y, sd.z, INFINITY);
            extend = true;
          }
        }
      }
    }
    uint32_t at = wave_append(a.shadow_count, shadow);
    if (shadow) a.shadow_queue[at] = p;
    uint32_t at2 = wave_append(out_count, extend);
    if (extend) out_queue[at2] = p;
  }
}

// Folds the batch's samples into SH9 per texel (baker.cpp:284-294) with
// Neumaier-compensated FP32 sums, in the reference order per sample: texel
// NEE term along its unnormalised direction, then 2*pi*Li along the primary
// direction.
__global__ void k_project(BakeArgs a) {
  const uint32_t T = a.texels_in_batch;
  uint32_t tl = blockIdx.x * blockDim.x + threadIdx.x;
  if (tl >= T) return;
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

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  for (; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_fill_f32(float* p, float v, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  for (; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ------------------------------------------------------------ launchers
template <bool ANY, bool COUNT, class Epi>
static hipError_t launch_trace_t(const TraceArgs& a, const Epi& e, int grid,
                                 int block, hipStream_t st) {
  size_t lds = (size_t)kStackLds * block * sizeof(uint2);
  hipLaunchKernelGGL((k_trace<ANY, COUNT, Epi>), dim3(grid), dim3(block), lds, st, a, e);
  return hipGetLastError();
}

hipError_t launch_trace_hits(const TraceArgs& a, float4* hits, bool count, int grid,
                             int block, hipStream_t st) {
  HitEpilogue e{hits};
  return count ? launch_trace_t<false, true>(a, e, grid, block, st)
               : launch_trace_t<false, false>(a, e, grid, block, st);
}

hipError_t launch_trace_occluded(const TraceArgs& a, uint32_t* occ, bool count,
                                 int grid, int block, hipStream_t st) {
  OccludedEpilogue e{occ};
  return count ? launch_trace_t<true, true>(a, e, grid, block, st)
               : launch_trace_t<true, false>(a, e, grid, block, st);
}

hipError_t launch_trace_bake_shadow(const TraceArgs& a, const BakeArgs& b, bool count,
                                    int grid, int block, hipStream_t st) {
  BakeShadowEpilogue e{b.li, b.scontrib, b.nee0, b.P};
  return count ? launch_trace_t<true, true>(a, e, grid, block, st)
               : launch_trace_t<true, false>(a, e, grid, block, st);
}

hipError_t launch_hit_details(const SceneView& sc, const float4* rays, const float4* hits,
                              uint32_t n, shb_hit* out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_hit_details, dim3((n + 255) / 256), dim3(256), 0, st, sc, rays, hits,
                     n, out);
  return hipGetLastError();
}

hipError_t launch_generate(const BakeArgs& b, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_generate, dim3(grid), dim3(256), 0, st, b);
  return hipGetLastError();
}

hipError_t launch_shade(const BakeArgs& b, int depth, const uint32_t* in_q,
                        const uint32_t* in_n, uint32_t* out_q, uint32_t* out_n, int grid,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_shade, dim3(grid), dim3(256), 0, st, b, depth, in_q, in_n, out_q, out_n);
  return hipGetLastError();
}

hipError_t launch_project(const BakeArgs& b, hipStream_t st) {
  uint32_t T = b.texels_in_batch;
  if (T == 0) return hipSuccess;
  hipLaunchKernelGGL(k_project, dim3((T + 127) / 128), dim3(128), 0, st, b);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  int grid = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}

hipError_t launch_fill_f32(float* p, float v, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  int grid = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_f32, dim3(grid), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}

int trace_blocks_per_cu(bool any, int block) {
  int nb = 0;
  size_t lds = (size_t)kStackLds * block * sizeof(uint2);
  hipError_t e;
  if (any)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &nb, reinterpret_cast<const void*>(&k_trace<true, false, BakeShadowEpilogue>), block,
        lds);
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &nb, reinterpret_cast<const void*>(&k_trace<false, false, HitEpilogue>), block, lds);
  if (e != hipSuccess || nb <= 0) nb = 1;
  return nb;
}

}  // namespace shb
