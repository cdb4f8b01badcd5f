// Wavefront path tracer for the SH9 lightmap bake on CDNA4.
//
// A batch holds T texels x spp samples as path slots p = sample * T + texel
// (sample-major, so the SH projection reads coalesced). Per bounce depth d:
//   k_trace<closest>  over the extension queue        (rtcIntersect1)
//   k_shade           hit -> emission, NEE pick, material sample, next ray
//   k_trace<any>      over the shadow queue, epilogue adds the NEE term
// k_generate makes the primary rays and the texel NEE shadow rays; k_project
// folds each texel's samples into 27 SH coefficients with compensated sums.
// Transmitted rays of alpha < 1 splits run in spare slots of the batch;
// k_merge_branches adds their radiance to their parents' before k_project.
// Mirrors BakeSHLightMap / Trace / EvaluateLightSamples /
// AccumulateIncomingLightSamples (reference src/baker.cpp:70-304,
// src/light.cpp:82-293).
#include <algorithm>

#include "../../include/shb.h"
#include "bake_device.h"
#include "kernels.h"
#include "traverse.h"
#include "bake_path.h"

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

// Appends the predicated lanes of a 256-thread block to NQ queues with one
// atomic per queue and block: a quarter of wave_append's atomics, which all
// land on one cache line of counters (the bake's shading stages made one
// to two per wave of every path). Queue q's entries of the block are
// contiguous, wave by wave and lane by lane. Every thread of the block must
// call it with the same NQ (it synchronises the block); `lds` holds 5 * NQ
// words. at[q]: the slot of a lane with pred[q].
template <int NQ>
__device__ __forceinline__ void block_append(uint32_t* const (&counter)[NQ],
                                             const bool (&pred)[NQ], uint32_t (&at)[NQ],
                                             uint32_t* lds) {
  // the wave index as a scalar, and lane prefixes counted from the ballot
  // (mbcnt), so nothing derived from the lane id is kept live across the
  // caller's code
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  uint64_t m[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    m[q] = __ballot(pred[q]);
    if (threadIdx.x == (w << 6)) lds[4 * q + w] = (uint32_t)__popcll(m[q]);
  }
  __syncthreads();
  if (threadIdx.x < (uint32_t)NQ) {
    const uint32_t q = threadIdx.x;
    uint32_t* c = counter[0];
#pragma unroll
    for (int k = 1; k < NQ; ++k)
      if (q == (uint32_t)k) c = counter[k];
    const uint32_t n0 = lds[4 * q], n1 = lds[4 * q + 1], n2 = lds[4 * q + 2], n3 = lds[4 * q + 3];
    lds[4 * q] = 0u;
    lds[4 * q + 1] = n0;
    lds[4 * q + 2] = n0 + n1;
    lds[4 * q + 3] = n0 + n1 + n2;
    const uint32_t total = n0 + n1 + n2 + n3;
    lds[4 * NQ + q] = total ? atomicAdd(c, total) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NQ; ++q)
    at[q] = lds[4 * NQ + q] + lds[4 * q + w] +
            __builtin_amdgcn_mbcnt_hi((uint32_t)(m[q] >> 32),
                                      __builtin_amdgcn_mbcnt_lo((uint32_t)m[q], 0u));
  __syncthreads();  // the words are rewritten by the next call
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

// Shadow rays of the bake: slot < Pc is a vertex NEE ray (adds its
// contribution to the path's radiance), slot >= Pc a texel NEE ray (marks
// the texel term visible).
struct BakeShadowEpilogue {
  float4* li;
  const float4* contrib;
  float4* nee0;
  uint32_t Pc;
  __device__ void operator()(uint32_t slot, const TravState& s) const {
    if (s.best != 0x7fffffff) return;
    if (slot < Pc) {
      float4 c = contrib[slot];
      float4 l = li[slot];
      l.x += c.x;
      l.y += c.y;
      l.z += c.z;
      li[slot] = l;
    } else {
      nee0[slot - Pc].w = 1.0f;
    }
  }
};

// Shadow rays of a bake with num_light_samples > 1: a visible pick sets its
// flag (vertex picks: scontrib.w at slot < vlimit = nls * Pc; texel picks:
// nee0.w at slot - vlimit); k_nee_gather sums the vertex picks in order.
struct BakeShadowMultiEpilogue {
  float4* scontrib;
  float4* nee0;
  uint32_t vlimit;
  __device__ void operator()(uint32_t slot, const TravState& s) const {
    if (s.best != 0x7fffffff) return;
    if (slot < vlimit)
      scontrib[slot].w = 1.0f;
    else
      nee0[slot - vlimit].w = 1.0f;
  }
};

// Shadow rays of the adjoint replay: a visible vertex NEE ray flags its slot
// (its gradient is added by k_adjoint_vertex), a visible texel NEE ray
// marks the texel term as in the forward bake.
struct BakeShadowAdjEpilogue {
  uint32_t* vflag;
  float4* nee0;
  uint32_t Pc;
  __device__ void operator()(uint32_t slot, const TravState& s) const {
    if (s.best != 0x7fffffff) return;
    if (slot < Pc)
      vflag[slot] = 1u;
    else
      nee0[slot - Pc].w = 1.0f;
  }
};

// Traversal stack in LDS: entry-major, lane-minor (conflict-free per wave).
struct LdsStack {
  uint2* mem;  // the kernel's dynamic __shared__ array
  int tid;
  int stride;
  int entries;
  // sp < 17 and stride <= 1024: a 24-bit multiply (full rate) suffices
  __device__ uint2 get(int sp) const { return mem[__umul24(sp, stride) + tid]; }
  __device__ void set(int sp, uint2 v) const { mem[__umul24(sp, stride) + tid] = v; }
  __device__ int cap() const { return entries; }
};

// Ray setup shared by the traversal kernels: the ray starts at its job's
// BVH root (a batch of several jobs traverses their concatenated trees).
// Any-hit rays visit each node's children far to near: the answer does not
// depend on the order, and the bake's shadow rays mostly end on geometry far
// along the ray (a point light's ray runs on (|d| - 0.001) |d| past the
// shading point, beyond the light, into the walls).
template <bool ANY>
__device__ __forceinline__ void trace_begin(const TraceArgs& a, uint32_t slot, TravRay& r,
                                            TravState& s) {
  const float4 o = a.rays[2 * (size_t)slot];
  const float4 d = a.rays[2 * (size_t)slot + 1];
  trav_setup(r, o, d);
  if (ANY) r.sstar ^= 7u;
  trav_init(s, r, d.w);
  uint32_t root = a.root0;
  if (a.nseg > 1) {
    // shadow slots: vertex pick j of slot p at j * Pc + p, texel pick j at
    // nls * Pc + j * P + p (one pick: p and Pc + p)
    const uint32_t nls = a.nls > 1 ? a.nls : 1u;
    const uint32_t vlimit = nls * a.Pc;
    uint32_t p = slot;
    if (p >= vlimit) {
      p -= vlimit;
      while (p >= a.P) p -= a.P;  // pick j: at most nls - 1 steps
    } else {
      while (p >= a.Pc) p -= a.Pc;
      if (p >= a.P) p = a.link[p - a.P].x;  // a branch slot: its root
    }
    root = a.map->root[find_segment(a.map->slot_base, a.nseg, p)];
  }
  s.gbase = root;
  // empty scene, or a ray without a usable direction / origin: miss
  if (!a.has_nodes || root == kNoRoot || !ray_is_finite(o, d)) s.gbits = 0;
}

// Error report, optional visit counts and the epilogue of a finished ray.
template <bool COUNT, class Epi>
__device__ __forceinline__ void trace_finish(const TraceArgs& a, const Epi& epi, uint32_t slot,
                                             uint32_t i, TravState& s, bool overflow,
                                             bool capped, uint32_t steps) {
  if (overflow || capped) {
    // error[0]: bit 0 stack overflow, bit 1 step cap; the first failing ray
    // records slot + 1, its stack depth and its step count in error[1..3]
    atomicOr(a.error, (overflow ? 1u : 0u) | (capped ? 2u : 0u));
    if (atomicCAS(a.error + 1, 0u, slot + 1u) == 0u) {
      a.error[2] = (uint32_t)s.sp;
      a.error[3] = steps;
    }
    s.best = 0x7fffffff;
  }
  if (COUNT) {
    if (a.ray_counts) {
      a.ray_counts[2 * (size_t)slot] = s.nodes;
      a.ray_counts[2 * (size_t)slot + 1] = s.tris;
    }
    if (a.totals) {
      unsigned long long* t = a.totals + 2 * (i & 63u);
      atomicAdd(t, (unsigned long long)s.nodes);
      atomicAdd(t + 1, (unsigned long long)s.tris);
    }
  }
  epi(slot, s);
}

// One ray per thread over a queue whose length is read on the device, so
// the grid is sized for the queue's capacity and surplus threads exit.
template <bool ANY, bool COUNT, class Epi>
__global__ void __launch_bounds__(256) k_trace(TraceArgs a, Epi epi) {
  extern __shared__ uint2 stack_mem[];  // a.stack_entries * blockDim.x
  const uint32_t count = a.count_ptr ? *a.count_ptr : a.count_const;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t slot = a.queue ? a.queue[i] : i;
  TravRay r;
  TravState s;
  trace_begin<ANY>(a, slot, r, s);
  LdsStack stack{stack_mem, (int)threadIdx.x, (int)blockDim.x, (int)a.stack_entries};
  bool overflow = false;
  bool capped = false;
  uint32_t steps = 0;
  while (!trav_step<ANY, COUNT>(a.scene, r, s, stack, overflow)) {
    // a ray visits each node at most once: more steps means a corrupt tree
    if (++steps > a.max_steps) {
      capped = true;
      break;
    }
  }
  trace_finish<COUNT>(a, epi, slot, i, s, overflow, capped, steps);
}

// Triangle tests of one traversal step spread over the wave. After every
// lane has decoded its node, the (owner lane, triangle) pairs of the whole
// wave go into a list in LDS (prefix sum of the per-lane counts by ballots)
// and each lane tests one pair per round, with the owner's ray fetched by
// cross-lane shuffles: about ceil(K / 64) rounds for K pairs instead of the
// largest per-lane count. Each owner's closest hit is the minimum of the
// 64-bit keys (t, id), the closest-hit rule itself, taken with LDS atomics;
// the winning lane then writes its t, u, v. The result equals the lane's own
// sequential loop (trav_tris): same accepted set, same minimum.
constexpr uint32_t kWavePairs = 256;  // K above this: each lane loops alone

struct WaveTriLds {
  float4 hit[64];                // owner's winning (t, u, v)
  unsigned long long key[64];    // owner's best key
  uint32_t pair[kWavePairs];     // owner lane << 26 | leaf-order triangle
};

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive prefix sum of x over the wave's 64 lanes: shifts by 1, 2, 4 and
// 8 lanes within each row of 16 (DPP row_shr; lanes shifted in from outside
// the row add 0), then the last lane of row 0 added to row 1 and of row 2 to
// row 3 (row_bcast:15), then lane 31 to rows 2 and 3 (row_bcast:31). Six
// DPP adds (VALU only, no memory access); each reads the register the
// previous instruction wrote, so two wait states (s_nop 1) sit before it.
// Every lane of the wave must be active (DPP reads inactive lanes' stale
// registers): wave_tris is reached by the whole wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  uint32_t y;
  asm volatile(
      "s_nop 1\n\t"
      "v_add_u32_dpp %0, %1, %1 row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "s_nop 1\n\t"
      "v_add_u32_dpp %0, %0, %0 row_shr:2 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "s_nop 1\n\t"
      "v_add_u32_dpp %0, %0, %0 row_shr:4 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "s_nop 1\n\t"
      "v_add_u32_dpp %0, %0, %0 row_shr:8 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "s_nop 1\n\t"
      "v_add_u32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_add_u32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
      "s_nop 1"
      : "=&v"(y)
      : "v"(x));
  return y;
}

// Every lane of the wave calls this (active: the lane decoded a node this
// step). Returns ANY && a hit was taken, as trav_tris does.
template <bool ANY, bool COUNT>
__device__ __forceinline__ bool wave_tris(const SceneView& sc, const TravRay& r, TravState& s,
                                          bool active, uint32_t tri_base, uint32_t tmask,
                                          WaveTriLds* L) {
  const uint32_t lane = lane_id();
  const uint32_t k = active ? (uint32_t)__builtin_popcount(tmask) : 0u;
  if (COUNT) s.tris += k;
  // off = the sum of k over the lanes below this one, K over the wave
  // (k <= 24: 8 leaf slots of at most 3 triangles)
  const uint32_t incl = wave_incl_scan(k);
  const uint32_t off = incl - k;
  const uint32_t K = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
  if (K == 0u) return false;
  if (K > kWavePairs) {
    bool f = false;
    if (active) f = trav_tris<ANY, false>(sc, r, s, tri_base, tmask);
    return f;
  }
  const unsigned long long init = hit_key(s.tbest, s.best);
  // pair = owner lane << 26 | triangle (tri_base + j < 2^26 when the wave
  // form is on)
  const uint32_t pbase = (lane << 26) + tri_base;
  uint32_t tm = k ? tmask : 0u, q = off;
  while (tm) {
    const uint32_t j = __builtin_ctz(tm);
    tm &= tm - 1u;
    L->pair[q++] = pbase + j;
  }
  if (k) L->key[lane] = init;
  wave_sync_lds();
  // the owner's origin in the component order of its major axis
  float rx, ry, rz;
  rot3(r.kz, r.ox, r.oy, r.oz, rx, ry, rz);
  for (uint32_t base = 0; base < K; base += 64u) {
    const uint32_t qi = base + lane;
    const bool have = qi < K;
    const uint32_t pr = have ? L->pair[qi] : 0u;
    const int o = (int)(pr >> 26);
    const float ox = __shfl(rx, o), oy = __shfl(ry, o), oz = __shfl(rz, o);
    const float Sx = __shfl(r.Sx, o), Sy = __shfl(r.Sy, o), Sz = __shfl(r.Sz, o);
    const int kz = __shfl(r.kz, o);
    const float tnear_o = __shfl(r.tnear, o);
    unsigned long long key = ~0ull;
    float t = 0.f, u = 0.f, v = 0.f;
    if (have) {
      // fewer than 2^24 triangle records when the wave form is on
      const float4* tp = reinterpret_cast<const float4*>(
          reinterpret_cast<const char*>(sc.tris) + __umul24(pr & 0x3ffffffu, 48u));
      const float4 a = tp[0], b = tp[1], c = tp[2];
      if (tri_test(kz, ox, oy, oz, Sx, Sy, Sz, a, b, c, t, u, v)) {
        const int id = (int)as_uint(a.w);
        // the owner's key starts at its best hit so far, so the minimum
        // over the keys keeps exactly the candidates better_hit accepts
        if (t >= tnear_o) {
          key = hit_key(t, id);
          atomicMin(&L->key[o], key);
        }
      }
    }
    wave_sync_lds();
    if (key != ~0ull && L->key[o] == key) L->hit[o] = make_float4(t, u, v, 0.f);
    wave_sync_lds();
  }
  bool found = false;
  if (k) {
    const unsigned long long fin = L->key[lane];
    if (fin != init) {
      const float4 h = L->hit[lane];
      s.tbest = h.x;
      s.best = (int)(uint32_t)fin;
      s.bu = h.y;
      s.bv = h.z;
      found = true;
    }
  }
  wave_sync_lds();  // the next step's list reuses the area
  return ANY && found;
}

// Persistent form: every wave keeps up to 64 rays in flight. When at least
// a.refill lanes are idle they take the next entries of the wave's pool, a
// range of kFetch queue positions claimed with one atomic on a.head, so lanes
// do not wait for the wave's longest ray and most refills need no atomic.
// Exit: once a claim reaches the queue's end the wave is `drained` (a.head
// only grows); it leaves when drained, its pool is empty and no lane holds a
// ray. Every ray ends by the step cap at the latest.
constexpr uint32_t kFetch = 128;  // >= 64: one claim serves every idle lane

template <bool ANY, bool COUNT, class Epi>
__global__ void __launch_bounds__(256, 6) k_trace_persistent(TraceArgs a, Epi epi) {
  // a.stack_entries * blockDim.x stack entries, then (with a.wave_tris) one
  // WaveTriLds per wave of the block
  extern __shared__ __align__(16) uint2 stack_mem[];
  WaveTriLds* tri_lds = reinterpret_cast<WaveTriLds*>(stack_mem + a.stack_entries * blockDim.x) +
                        threadIdx.x / 64u;
  const uint32_t count = a.count_ptr ? *a.count_ptr : a.count_const;
  const uint32_t lane = lane_id();
  const uint64_t lanes_below = (1ull << lane) - 1ull;
  LdsStack stack{stack_mem, (int)threadIdx.x, (int)blockDim.x, (int)a.stack_entries};
  TravRay r;
  TravState s;
  uint32_t slot = 0, qi = 0, steps = 0;
  bool has = false, overflow = false;
  // wave-uniform pool [pool_next, pool_end) of claimed queue positions
  uint32_t pool_next = 0, pool_end = 0;
  bool drained = false;
  while (true) {
    const uint64_t idle = __ballot(!has);
    const uint32_t n_idle = (uint32_t)__popcll(idle);
    const bool pool_empty = pool_next >= pool_end;
    if (n_idle >= a.refill && !(drained && pool_empty)) {
      const uint32_t take_old = min(pool_end - pool_next, n_idle);
      const uint32_t old_base = pool_next;
      pool_next += take_old;
      uint32_t new_base = 0;
      bool fresh = false;  // a claim of this pass holds positions < count
      if (take_old < n_idle && !drained) {
        const int leader = __ffsll((unsigned long long)idle) - 1;
        if ((int)lane == leader) new_base = atomicAdd(a.head, kFetch);
        new_base = __shfl(new_base, leader);
        if (new_base >= count) {
          drained = true;
          pool_next = pool_end = 0;
        } else {
          if (new_base + kFetch >= count) drained = true;
          pool_end = min(new_base + kFetch, count);
          pool_next = min(new_base + (n_idle - take_old), pool_end);
          fresh = true;
        }
      }
      if (!has) {
        const uint32_t rank = (uint32_t)__popcll(idle & lanes_below);
        bool got = false;
        if (rank < take_old) {
          qi = old_base + rank;
          got = true;
        } else if (fresh) {
          qi = new_base + (rank - take_old);
          got = qi < count;
        }
        if (got) {
          slot = a.queue ? a.queue[qi] : qi;
          trace_begin<ANY>(a, slot, r, s);
          has = true;
          overflow = false;
          steps = 0;
        }
      }
    }
    if (__ballot(has) == 0ull) {
      if (drained && pool_next >= pool_end) break;
      continue;  // all idle, more work claimable: the next pass refills
    }
    if (a.wave_tris) {
      // node half per lane, then the wave's triangle tests together
      uint32_t tb = 0u, tmk = 0u;
      bool done = false;
      if (has) done = trav_node<COUNT>(a.scene, r, s, stack, overflow, tb, tmk);
      const bool found = wave_tris<ANY, COUNT>(a.scene, r, s, has && !done, tb, tmk, tri_lds);
      if (has) {
        if (found) done = true;
        bool capped = false;
        if (!done && ++steps > a.max_steps) {
          capped = true;
          done = true;
        }
        if (done) {
          trace_finish<COUNT>(a, epi, slot, qi, s, overflow, capped, steps);
          has = false;
        }
      }
    } else if (has) {
      bool done = trav_step<ANY, COUNT>(a.scene, r, s, stack, overflow);
      bool capped = false;
      if (!done && ++steps > a.max_steps) {
        capped = true;
        done = true;
      }
      if (done) {
        trace_finish<COUNT>(a, epi, slot, qi, s, overflow, capped, steps);
        has = false;
      }
    }
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
// Primary ray + texel NEE for every path slot of the batch (UNIT: every
// light and albedo scale is 1).
template <bool UNIT, bool MULTI, bool ONE>
__global__ void __launch_bounds__(256, 5) k_generate(BakeArgs a) {
  __shared__ uint32_t qw[10];
  const uint32_t n = a.P;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    uint32_t p = base + threadIdx.x;
    uint32_t shadow = 0u;
    if (p < n) {
      const SlotRef sr = slot_ref(a, p);
      with_job<ONE>(a, sr.job,
                    [&](const JobDev& J) { shadow = generate_path<UNIT, MULTI>(a, J, sr, p); });
    }
    // shadow rays, and the shadow queries answered without one (counted)
    uint32_t* const sq[2] = {a.shadow_count, a.resolved_count};
    if (MULTI) {
      for (uint32_t j = 0; j < a.nls; ++j) {  // block-uniform bound
        const bool made[2] = {((shadow >> j) & 1u) != 0u, ((shadow >> (16 + j)) & 1u) != 0u};
        uint32_t at[2];
        block_append<2>(sq, made, at, qw);
        if (made[0]) a.shadow_queue[at[0]] = a.nls * a.Pc + j * a.P + p;
      }
    } else {
      const bool made[2] = {(shadow & 0xffffu) != 0u, (shadow >> 16) != 0u};
      uint32_t at[2];
      block_append<2>(sq, made, at, qw);
      if (made[0]) a.shadow_queue[at[0]] = a.Pc + p;
    }
  }
}

// Hit shading at depth d (ADJ: the replay form of the adjoint, with the
// block's gradient sums in dynamic LDS; SPLIT: batches with spare slots for
// alpha splits; UNIT: every light and albedo scale is 1; ONE: the batch is
// one job's segment, so the shading code is compiled once, for a
// wave-uniform job record).
// Every form is capped at 96 registers (at least 5 waves per SIMD), since
// the light loop waits on a scalar load per light: the bake's one-job
// forms take 75-81 (6 waves), the other one-job forms up to 91; the
// multi-job forms, with a second copy of the shading code for waves that
// hold several jobs, spill a little under the cap, which is still faster
// than 4 spill-free waves.
template <bool ADJ, bool SPLIT, bool UNIT, bool MULTI, bool ONE>
__global__ void __launch_bounds__(256, (UNIT && ONE && !MULTI && !SPLIT && !ADJ) ? 6 : 5) k_shade(BakeArgs a, int depth, const uint32_t* in_queue,
                                                  const uint32_t* in_count, uint32_t* out_queue,
                                                  uint32_t* out_count) {
  extern __shared__ float grad_lds[];
  __shared__ uint32_t qw[20];
  GradAcc acc;
  if (ADJ)
    acc.init(a.adj_nparam <= kAdjLdsParams ? grad_lds : nullptr, a.adj_grad, a.adj_nparam);
  const uint32_t n = in_queue ? *in_count : a.P;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    uint32_t i = base + threadIdx.x;
    uint32_t shadow = 0u;
    bool extend = false;
    uint32_t p = 0, child = kNoSlot;
    if (i < n) {
      p = in_queue ? in_queue[i] : i;
      const SlotRef sr = slot_ref(a, SPLIT ? root_slot(a, p) : p);
      with_job<ONE>(a, sr.job, [&](const JobDev& J) {
        shade_path<ADJ, SPLIT, UNIT, MULTI>(a, J, sr, depth, p, shadow, extend, child, &acc);
        // the stored shadow ray, read back with its light (scontrib.w) now
        // that the path's values are dead, answered by the emitter's triangle
        // test where it can be (single-triangle emitters: the sampled
        // triangle is not stored)
        if (!MULTI && shadow == 1u) {
          const float4 d = a.sray[2 * (size_t)p + 1];
          if (d.w >= 1.01f) {
            const LightDev& L = J.scene.lights[as_uint(a.scontrib[p].w)];
            if (L.num_tris == 1 && emitter_occludes(J.scene, L, 0u, a.sray[2 * (size_t)p], d))
              shadow = 1u << 16;
          }
        }
      });
    }
    // shadow rays (one queue per pick j when MULTI), the continuation and,
    // with SPLIT, the transmitted ray of a split (the same queue); shadow
    // queries answered without a ray are counted
    if (MULTI) {
      uint32_t* const sq[2] = {a.shadow_count, a.resolved_count};
      for (uint32_t j = 0; j < a.nls; ++j) {  // block-uniform bound
        const bool made[2] = {((shadow >> j) & 1u) != 0u, ((shadow >> (16 + j)) & 1u) != 0u};
        uint32_t at[2];
        block_append<2>(sq, made, at, qw);
        if (made[0]) a.shadow_queue[at[0]] = j * a.Pc + p;
      }
    }
    const bool made = !MULTI && (shadow & 0xffffu) != 0u;
    const bool resolved = !MULTI && (shadow >> 16) != 0u;
    if (SPLIT) {
      uint32_t* const qs[4] = {a.shadow_count, out_count, out_count, a.resolved_count};
      const bool pr[4] = {made, extend, child != kNoSlot, resolved};
      uint32_t at[4];
      block_append<4>(qs, pr, at, qw);
      if (pr[0]) a.shadow_queue[at[0]] = p;
      if (pr[1]) out_queue[at[1]] = p;
      if (pr[2]) out_queue[at[2]] = child;
    } else {
      uint32_t* const qs[3] = {a.shadow_count, out_count, a.resolved_count};
      const bool pr[3] = {made, extend, resolved};
      uint32_t at[3];
      block_append<3>(qs, pr, at, qw);
      if (pr[0]) a.shadow_queue[at[0]] = p;
      if (pr[1]) out_queue[at[1]] = p;
    }
  }
  if (ADJ) acc.flush();
}

// num_light_samples > 1: each slot k_shade took at this depth adds its
// visible vertex picks, in pick order, divided by its job's sample count
// (EvaluateLightSamples, light.cpp:164-188: sum over the samples / n).
__global__ void k_nee_gather(BakeArgs a, const uint32_t* in_queue, const uint32_t* in_count) {
  const uint32_t n = in_queue ? *in_count : a.P;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t p = in_queue ? in_queue[i] : i;
    f3 sum = {0.f, 0.f, 0.f};
    bool any = false;
    for (uint32_t j = 0; j < a.nls; ++j) {
      const float4 c = a.scontrib[(size_t)j * a.Pc + p];
      if (c.w != 0.f) {
        sum = sum + f3_of(c);
        any = true;
      }
    }
    if (!any) continue;
    const SlotRef sr = slot_ref(a, root_slot(a, p));
    const float nj = (float)a.jobs[sr.job].nls;
    float4 l = a.li[p];
    l.x += sum.x / nj;
    l.y += sum.y / nj;
    l.z += sum.z / nj;
    a.li[p] = l;
  }
}

// Adjoint of the visible vertex NEE terms of one depth: the shadow queue's
// entries below Pc (texel NEE rays sit at Pc + slot).
__global__ void k_adjoint_vertex(BakeArgs a, int depth, const uint32_t* sh_count) {
  extern __shared__ float grad_lds[];
  GradAcc acc;
  acc.init(a.adj_nparam <= kAdjLdsParams ? grad_lds : nullptr, a.adj_grad, a.adj_nparam);
  const uint32_t n = *sh_count;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t slot = a.shadow_queue[i];
    if (slot >= a.Pc || a.adj_vflag[slot] == 0u) continue;
    a.adj_vflag[slot] = 0u;
    const SlotRef sr = slot_ref(a, root_slot(a, slot));
    adjoint_vertex(a, a.jobs[sr.job], depth, slot, &acc);
  }
  acc.flush();
}

// One thread per texel of the batch. The ADJ form ends with one block-wide
// flush that every thread reaches (a single barrier per wave).
template <bool ADJ>
__global__ void k_project(BakeArgs a) {
  extern __shared__ float grad_lds[];
  GradAcc acc;
  if (ADJ)
    acc.init(a.adj_nparam <= kAdjLdsParams ? grad_lds : nullptr, a.adj_grad, a.adj_nparam);
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < a.T) {
    uint32_t job, rank, slot0, stride;
    if (a.nseg <= 1) {
      job = a.job0;
      rank = a.texel_begin0 + t;
      slot0 = t;
      stride = a.T0;
    } else {
      const BatchMap& m = *a.map;
      const uint32_t s = find_segment(m.tex_base, a.nseg, t);
      const uint32_t tl = t - m.tex_base[s];
      job = m.job[s];
      rank = m.texel_begin[s] + tl;
      slot0 = m.slot_base[s] + tl;
      stride = m.tex_base[s + 1] - m.tex_base[s];
    }
    if (ADJ)
      project_texel_adjoint(a, a.jobs[job], rank, slot0, stride, &acc);
    else
      with_job<false>(a, job, [&](const JobDev& J) { project_texel(a, J, rank, slot0, stride); });
  }
  if (ADJ) acc.flush();
}

// Branch slots spawned at `depth` add their radiance to their parents'. A
// path shades one vertex per depth, so no parent gains two branches in one
// launch, and a branch spawned at `depth` spawns its own only deeper, so it
// is complete (merged into by the earlier, deeper launches) when read here:
// the sums do not depend on the order of the slots.
__global__ void k_merge_branches(BakeArgs a, int depth) {
  const uint32_t n = min(*a.spawn_count, a.S);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint4 L = a.link[i];
    if ((int)L.z != depth) continue;
    const float4 c = a.li[a.P + i];
    float4 l = a.li[L.y];
    l.x += c.x;
    l.y += c.y;
    l.z += c.z;
    a.li[L.y] = l;
  }
}

__global__ void k_relocate_nodes(const CWNode* src, CWNode* dst, uint32_t n, uint32_t node_off,
                                 uint32_t tri_off) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // both arrays at the device node stride
  CWNode nd = *reinterpret_cast<const CWNode*>(reinterpret_cast<const char*>(src) +
                                               (size_t)i * kNodeBytes);
  nd.child_base += node_off;
  nd.tri_base += tri_off;
  *reinterpret_cast<CWNode*>(reinterpret_cast<char*>(dst) + (size_t)i * kNodeBytes) = nd;
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
static hipError_t launch_trace_t(const TraceArgs& a, const Epi& e, uint32_t capacity, int block,
                                 hipStream_t st) {
  if (capacity == 0) return hipSuccess;
  if (a.stack_entries < 1 || a.stack_entries > (uint32_t)kStackLds) return hipErrorInvalidValue;
  size_t lds = (size_t)a.stack_entries * block * sizeof(uint2);
  uint32_t grid = (capacity + block - 1) / block;
  if (a.head && a.refill >= 1 && a.refill <= 64) {
    if (a.wave_tris) lds += (size_t)(block / 64) * sizeof(WaveTriLds);
    // persistent: as many blocks as fit on the device at once (occupancy of
    // this kernel with its LDS stack), no more than the queue could use
    int per_cu = 0;
    auto kern = k_trace_persistent<ANY, COUNT, Epi>;
    hipError_t oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, lds);
    if (oe != hipSuccess || per_cu < 1) per_cu = 1;
    uint32_t resident = (uint32_t)std::max(1, a.num_cus) * (uint32_t)per_cu;
    grid = std::min(grid, resident);
    hipLaunchKernelGGL((k_trace_persistent<ANY, COUNT, Epi>), dim3(grid), dim3(block), lds, st,
                       a, e);
  } else {
    hipLaunchKernelGGL((k_trace<ANY, COUNT, Epi>), dim3(grid), dim3(block), lds, st, a, e);
  }
  return hipGetLastError();
}

hipError_t launch_trace_hits(const TraceArgs& a, float4* hits, bool count, uint32_t capacity,
                             int block, hipStream_t st) {
  HitEpilogue e{hits};
  return count ? launch_trace_t<false, true>(a, e, capacity, block, st)
               : launch_trace_t<false, false>(a, e, capacity, block, st);
}

hipError_t launch_trace_occluded(const TraceArgs& a, uint32_t* occ, bool count,
                                 uint32_t capacity, int block, hipStream_t st) {
  OccludedEpilogue e{occ};
  return count ? launch_trace_t<true, true>(a, e, capacity, block, st)
               : launch_trace_t<true, false>(a, e, capacity, block, st);
}

hipError_t launch_trace_bake_shadow(const TraceArgs& a, const BakeArgs& b, bool count,
                                    uint32_t capacity, int block, hipStream_t st) {
  BakeShadowEpilogue e{b.li, b.scontrib, b.nee0, b.Pc};
  return count ? launch_trace_t<true, true>(a, e, capacity, block, st)
               : launch_trace_t<true, false>(a, e, capacity, block, st);
}

hipError_t launch_trace_bake_shadow_multi(const TraceArgs& a, const BakeArgs& b, bool count,
                                          uint32_t capacity, int block, hipStream_t st) {
  BakeShadowMultiEpilogue e{b.scontrib, b.nee0, b.nls * b.Pc};
  return count ? launch_trace_t<true, true>(a, e, capacity, block, st)
               : launch_trace_t<true, false>(a, e, capacity, block, st);
}

hipError_t launch_nee_gather(const BakeArgs& b, const uint32_t* in_q, const uint32_t* in_n,
                             int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_nee_gather, dim3(grid), dim3(256), 0, st, b, in_q, in_n);
  return hipGetLastError();
}

hipError_t launch_trace_bake_shadow_adj(const TraceArgs& a, const BakeArgs& b, bool count,
                                        uint32_t capacity, int block, hipStream_t st) {
  BakeShadowAdjEpilogue e{b.adj_vflag, b.nee0, b.Pc};
  return count ? launch_trace_t<true, true>(a, e, capacity, block, st)
               : launch_trace_t<true, false>(a, e, capacity, block, st);
}

hipError_t launch_hit_details(const SceneView& sc, const float4* rays, const float4* hits,
                              uint32_t n, shb_hit* out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_hit_details, dim3((n + 255) / 256), dim3(256), 0, st, sc, rays, hits,
                     n, out);
  return hipGetLastError();
}

hipError_t launch_generate(const BakeArgs& b, int grid, hipStream_t st) {
#define SHB_GEN(U, M, O) hipLaunchKernelGGL((k_generate<U, M, O>), dim3(grid), dim3(256), 0, st, b)
  const int v = (b.unit_scales ? 4 : 0) | (b.nls > 1 ? 2 : 0) | (b.nseg <= 1 ? 1 : 0);
  switch (v) {
    case 0: SHB_GEN(false, false, false); break;
    case 1: SHB_GEN(false, false, true); break;
    case 2: SHB_GEN(false, true, false); break;
    case 3: SHB_GEN(false, true, true); break;
    case 4: SHB_GEN(true, false, false); break;
    case 5: SHB_GEN(true, false, true); break;
    case 6: SHB_GEN(true, true, false); break;
    default: SHB_GEN(true, true, true); break;
  }
#undef SHB_GEN
  return hipGetLastError();
}

// Dynamic LDS of an adjoint kernel: the block's gradient sums, when they fit.
static size_t adj_lds(const BakeArgs& b) {
  return b.adj_nparam <= kAdjLdsParams ? (size_t)b.adj_nparam * sizeof(float) : 0;
}

hipError_t launch_shade(const BakeArgs& b, int depth, const uint32_t* in_q,
                        const uint32_t* in_n, uint32_t* out_q, uint32_t* out_n, int grid,
                        hipStream_t st) {
  const size_t lds = b.adj_grad ? adj_lds(b) : 0;
#define SHB_SHADE(A, S, U, M, O)                                                            \
  hipLaunchKernelGGL((k_shade<A, S, U, M, O>), dim3(grid), dim3(256), lds, st, b, depth, in_q, \
                     in_n, out_q, out_n)
  // the adjoint runs one job (batches of one segment)
  if (b.adj_grad && b.nseg > 1) return hipErrorInvalidValue;
  const int v = (b.S ? 2 : 0) | (b.unit_scales ? 1 : 0);
  if (b.adj_grad) {
    switch (v) {
      case 0: SHB_SHADE(true, false, false, false, true); break;
      case 1: SHB_SHADE(true, false, true, false, true); break;
      case 2: SHB_SHADE(true, true, false, false, true); break;
      default: SHB_SHADE(true, true, true, false, true); break;
    }
    return hipGetLastError();
  }
  const int w = v | (b.nls > 1 ? 4 : 0) | (b.nseg <= 1 ? 8 : 0);
  switch (w) {
    case 0: SHB_SHADE(false, false, false, false, false); break;
    case 1: SHB_SHADE(false, false, true, false, false); break;
    case 2: SHB_SHADE(false, true, false, false, false); break;
    case 3: SHB_SHADE(false, true, true, false, false); break;
    case 4: SHB_SHADE(false, false, false, true, false); break;
    case 5: SHB_SHADE(false, false, true, true, false); break;
    case 6: SHB_SHADE(false, true, false, true, false); break;
    case 7: SHB_SHADE(false, true, true, true, false); break;
    case 8: SHB_SHADE(false, false, false, false, true); break;
    case 9: SHB_SHADE(false, false, true, false, true); break;
    case 10: SHB_SHADE(false, true, false, false, true); break;
    case 11: SHB_SHADE(false, true, true, false, true); break;
    case 12: SHB_SHADE(false, false, false, true, true); break;
    case 13: SHB_SHADE(false, false, true, true, true); break;
    case 14: SHB_SHADE(false, true, false, true, true); break;
    default: SHB_SHADE(false, true, true, true, true); break;
  }
#undef SHB_SHADE
  return hipGetLastError();
}

hipError_t launch_adjoint_vertex(const BakeArgs& b, int depth, const uint32_t* sh_count, int grid,
                                 hipStream_t st) {
  hipLaunchKernelGGL(k_adjoint_vertex, dim3(grid), dim3(256), adj_lds(b), st, b, depth, sh_count);
  return hipGetLastError();
}

hipError_t launch_relocate_nodes(const CWNode* src, CWNode* dst, uint32_t n, uint32_t node_off,
                                 uint32_t tri_off, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_relocate_nodes, dim3((n + 255) / 256), dim3(256), 0, st, src, dst, n,
                     node_off, tri_off);
  return hipGetLastError();
}

hipError_t launch_project(const BakeArgs& b, hipStream_t st) {
  uint32_t T = b.T;
  if (T == 0) return hipSuccess;
  if (b.adj_grad)
    hipLaunchKernelGGL(k_project<true>, dim3((T + 127) / 128), dim3(128), adj_lds(b), st, b);
  else
    hipLaunchKernelGGL(k_project<false>, dim3((T + 127) / 128), dim3(128), 0, st, b);
  return hipGetLastError();
}

hipError_t launch_merge_branches(const BakeArgs& b, int depth, int grid, hipStream_t st) {
  if (b.S == 0) return hipSuccess;
  hipLaunchKernelGGL(k_merge_branches, dim3(grid), dim3(256), 0, st, b, depth);
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

}  // namespace shb
