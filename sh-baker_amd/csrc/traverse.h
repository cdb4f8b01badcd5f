// Ray traversal of the compressed 8-wide BVH on CDNA4 (wave64).
//
// Per lane: one ray, a current "node group" (child base index + 8-bit hit
// mask in octant order + the parent's internal-child mask) and a stack of
// up to kStackLds groups in LDS (laid out entry-major, lane-minor, so a
// wave's accesses hit 64 distinct banks); the group stack never holds more
// than one entry per tree level, and the upload rejects deeper trees.
// A step decodes one node (trav_node) and then tests the triangles of its
// leaf children before the next step: per lane (trav_tris, trav_step) or,
// in the persistent kernel, spread over the wave. Ray/triangle: the
// watertight test of Woop, Benthin, Wald (JCGT 2013), double sided, ties in
// t resolved to the lowest triangle id so the closest hit does not depend
// on traversal order.
#ifndef SHB_TRAVERSE_H_
#define SHB_TRAVERSE_H_

#include "shb_common.h"

namespace shb {

struct TravRay {
  float ox, oy, oz;
  float dx, dy, dz;
  float ix, iy, iz;     // 1/d with zero components replaced by +-1e20
  float tnear;
  int kz;               // major axis of d; kx = kz + 1, ky = kz + 2 (mod 3)
  float Sx, Sy, Sz;     // d[kx] / d[kz], d[ky] / d[kz], 1 / d[kz]
  uint32_t sstar;       // nearest octant slot
  uint32_t negx, negy, negz;
  uint32_t zx, zy, zz;  // direction component below 1e-20 in magnitude
};

// 1.0f / d, correctly rounded: the corrected hardware reciprocal where it
// equals the IEEE division bit for bit (|d| in [2^-48, 2^63], checked for
// every float of that range; the sign is copied, as the division's is),
// the division elsewhere.
__device__ __forceinline__ float recip(float d) {
  const float a = fabsf(d);
  if (a >= 0x1p-48f && a <= 0x1p63f) return copysignf(rcp_core(a), d);
  return 1.0f / d;
}

__device__ __forceinline__ float safe_inv(float d) {
  float ad = fabsf(d);
  float r = recip(d);
  return ad < 1e-20f ? (d < 0.f || (d == 0.f && signbit(d)) ? -1e20f : 1e20f) : r;
}

// Finite origin and a finite, non-zero direction (anything else cannot hit).
__device__ __forceinline__ bool ray_is_finite(float4 o, float4 d) {
  float sum = o.x + o.y + o.z + d.x + d.y + d.z;
  float dd = d.x * d.x + d.y * d.y + d.z * d.z;
  return (sum - sum) == 0.0f && dd > 0.0f && (dd - dd) == 0.0f;
}

__device__ __forceinline__ float sel3(int k, float x, float y, float z) {
  return k == 0 ? x : (k == 1 ? y : z);
}

__device__ __forceinline__ void trav_setup(TravRay& r, float4 o, float4 d) {
  r.ox = o.x;
  r.oy = o.y;
  r.oz = o.z;
  r.tnear = o.w;
  r.dx = d.x;
  r.dy = d.y;
  r.dz = d.z;
  r.ix = safe_inv(d.x);
  r.iy = safe_inv(d.y);
  r.iz = safe_inv(d.z);
  r.zx = fabsf(d.x) < 1e-20f;
  r.zy = fabsf(d.y) < 1e-20f;
  r.zz = fabsf(d.z) < 1e-20f;
  r.negx = d.x < 0.f;
  r.negy = d.y < 0.f;
  r.negz = d.z < 0.f;
  uint32_t oct = r.negx | (r.negy << 1) | (r.negz << 2);
  r.sstar = oct ^ 7u;
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  float dkz = sel3(kz, d.x, d.y, d.z);
  r.kz = kz;
  r.Sx = sel3(kx, d.x, d.y, d.z) / dkz;
  r.Sy = sel3(ky, d.x, d.y, d.z) / dkz;
  r.Sz = recip(dkz);
}

// The components (v[kx], v[ky], v[kz]) of v for major axis kz.
__device__ __forceinline__ void rot3(int kz, float x, float y, float z, float& rx, float& ry,
                                     float& rz) {
  rx = kz == 0 ? y : (kz == 1 ? z : x);
  ry = kz == 0 ? z : (kz == 1 ? x : y);
  rz = kz == 0 ? x : (kz == 1 ? y : z);
}

// Watertight test (see the host restatement for the identical arithmetic).
// (ox, oy, oz): the ray origin in the order (kx, ky, kz) = (kz + 1, kz + 2,
// kz) mod 3. The paper swaps kx and ky when d[kz] < 0 to keep the winding
// for backface culling; this test is double sided, and the swap negates U,
// V, W, det and T exactly (every operation involved is sign symmetric under
// round-to-nearest), so t, u and v -- and which triangles pass -- are the
// same bits with or without it; the host restatement keeps the swap.
__device__ __forceinline__ bool tri_test(int kz, float ox, float oy, float oz, float Sx,
                                         float Sy, float Sz, float4 a0, float4 b0, float4 c0,
                                         float& t, float& u, float& v) {
  float4 a, b, c;
  rot3(kz, a0.x, a0.y, a0.z, a.x, a.y, a.z);
  rot3(kz, b0.x, b0.y, b0.z, b.x, b.y, b.z);
  rot3(kz, c0.x, c0.y, c0.z, c.x, c.y, c.z);
  const float Akz = a.z - oz, Bkz = b.z - oz, Ckz = c.z - oz;
  float Ax = fmaf(-Sx, Akz, a.x - ox);
  float Ay = fmaf(-Sy, Akz, a.y - oy);
  float Bx = fmaf(-Sx, Bkz, b.x - ox);
  float By = fmaf(-Sy, Bkz, b.y - oy);
  float Cx = fmaf(-Sx, Ckz, c.x - ox);
  float Cy = fmaf(-Sy, Ckz, c.y - oy);
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.f || V == 0.f || W == 0.f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
  float det = U + V + W;
  if (det == 0.f) return false;
  float Az = Sz * Akz, Bz = Sz * Bkz, Cz = Sz * Ckz;
  float T = fmaf(U, Az, fmaf(V, Bz, W * Cz));
  float inv = recip(det);
  t = T * inv;
  u = V * inv;
  v = W * inv;
  return true;
}

struct TravState {
  uint32_t gbase;   // group child base
  uint32_t gbits;   // hits (bits 0-7, octant order) | imask << 8
  int sp;
  float tbest;
  int best;         // INT_MAX: none
  float bu, bv;
  uint32_t nodes, tris;  // visit counters
};

__device__ __forceinline__ void trav_init(TravState& s, const TravRay& r, float tfar) {
  s.gbase = 0;
  s.gbits = (1u << r.sstar) | (1u << 8);
  s.sp = 0;
  s.tbest = tfar;
  s.best = 0x7fffffff;
  s.bu = s.bv = 0.f;
  s.nodes = s.tris = 0;
}

// Order-preserving 64-bit key of a candidate hit: (t, id) compares as the
// closest-hit rule does (smaller t first, ties to the lower id). t + 0 maps
// -0 to +0 so equal t give equal keys.
__device__ __forceinline__ unsigned long long hit_key(float t, int id) {
  uint32_t u = as_uint(t + 0.0f);
  u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  return ((unsigned long long)u << 32) | (uint32_t)id;
}
__device__ __forceinline__ float key_t(unsigned long long k) {
  uint32_t u = (uint32_t)(k >> 32);
  u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
  return as_float(u);
}

// Candidate acceptance of the closest-hit rule (t in [tnear, tbest], ties in
// t to the lower id).
__device__ __forceinline__ bool better_hit(float t, int id, float tnear, float tbest, int best) {
  return t >= tnear && (t < tbest || (t == tbest && id < best));
}

// First half of a traversal step: pop a group when the current one is
// empty, visit its nearest child node and decode it; the node's child group
// becomes current and the triangles to test are returned as (tri_base,
// tmask). Returns true when the ray has nothing left to visit (tmask 0).
// Stack: any type with get(int) / set(int, uint2) / cap() (entries; at most
// kStackLds) — on the GPU a slice of the kernel's dynamic LDS, sized to the
// uploaded tree's depth, on the host a local array.
template <bool COUNT, class Stack>
__device__ __forceinline__ bool trav_node(const SceneView& sc, const TravRay& r, TravState& s,
                                          Stack& stack, bool& overflow, uint32_t& tri_base,
                                          uint32_t& tmask_out) {
  tmask_out = 0u;
  tri_base = 0u;
  if ((s.gbits & 0xffu) == 0u) {
    if (s.sp == 0) return true;
    --s.sp;
    uint2 e = stack.get(s.sp);
    s.gbase = e.x;
    s.gbits = e.y;
  }
  uint32_t hk = s.gbits & 0xffu;
  uint32_t k = __builtin_ctz(hk);
  s.gbits &= ~(1u << k);
  uint32_t slot = k ^ r.sstar;
  uint32_t im = (s.gbits >> 8) & 0xffu;
  uint32_t ni = s.gbase + __builtin_popcount(im & ((1u << slot) - 1u));
  if ((s.gbits & 0xffu) != 0u) {
    if (s.sp >= stack.cap()) {  // deeper than the BVH the upload accepted
      overflow = true;
      return true;
    }
    stack.set(s.sp, make_uint2(s.gbase, s.gbits));
    ++s.sp;
  }
  // ---- decode node ni
  // a 32-bit byte offset from one 24-bit multiply: trees hold fewer than
  // 2^24 nodes (kMaxTreeNodes, checked where trees are uploaded and joined)
  const uint4* np = reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(sc.nodes) +
                                                   __umul24(ni, kNodeBytes));
  uint4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3], w4 = np[4];
  if (COUNT) ++s.nodes;
  float px = as_float(w0.x), py = as_float(w0.y), pz = as_float(w0.z);
  uint32_t ew = w0.w;
  uint32_t imask = ew >> 24;
  // a = 2^e / d: the exact product of the node's power-of-two scale with 1/d
  float ax = ldexpf(r.ix, (int)(int8_t)(ew & 0xffu)), bx = (px - r.ox) * r.ix;
  float ay = ldexpf(r.iy, (int)(int8_t)((ew >> 8) & 0xffu)), by = (py - r.oy) * r.iy;
  float az = ldexpf(r.iz, (int)(int8_t)((ew >> 16) & 0xffu)), bz = (pz - r.oz) * r.iz;
  // Per-axis absolute bound on the rounding of fmaf(q, a, b) against the
  // exact plane distance (p + q s - o) / d: 16 ulps of (|b| + 256 |a|).
  // Near planes move down and far planes up by it, so no box holding a hit
  // in [tnear, tbest] is culled and closest-hit ties resolve as in a
  // brute-force search. Axes the ray is parallel to (d = 0, 1/d replaced by
  // +-1e20) keep an exact sign test. Computed as |a| 2^-12 + |b| 2^-20 in
  // one fma: the power-of-two scalings are exact, so this is the same float
  // as (|b| + 256 |a|) 2^-20.
  float ex = r.zx ? 0.f : fmaf(fabsf(ax), 0x1p-12f, fabsf(bx) * 0x1p-20f);
  float ey = r.zy ? 0.f : fmaf(fabsf(ay), 0x1p-12f, fabsf(by) * 0x1p-20f);
  float ez = r.zz ? 0.f : fmaf(fabsf(az), 0x1p-12f, fabsf(bz) * 0x1p-20f);
  const float bxn = bx - ex, bxf = bx + ex;
  const float byn = by - ey, byf = by + ey;
  const float bzn = bz - ez, bzf = bz + ez;
  // near / far plane words per axis by direction sign
  uint32_t nx0 = r.negx ? w3.z : w2.x, nx1 = r.negx ? w3.w : w2.y;
  uint32_t fx0 = r.negx ? w2.x : w3.z, fx1 = r.negx ? w2.y : w3.w;
  uint32_t ny0 = r.negy ? w4.x : w2.z, ny1 = r.negy ? w4.y : w2.w;
  uint32_t fy0 = r.negy ? w2.z : w4.x, fy1 = r.negy ? w2.w : w4.y;
  uint32_t nz0 = r.negz ? w4.z : w3.x, nz1 = r.negz ? w4.w : w3.y;
  uint32_t fz0 = r.negz ? w3.x : w4.z, fz1 = r.negz ? w3.y : w4.w;
  uint32_t hits = 0, tmask = 0;
  const float tmax_ray = s.tbest;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int sh = 8 * (c & 3);
    uint32_t nxw = c < 4 ? nx0 : nx1, fxw = c < 4 ? fx0 : fx1;
    uint32_t nyw = c < 4 ? ny0 : ny1, fyw = c < 4 ? fy0 : fy1;
    uint32_t nzw = c < 4 ? nz0 : nz1, fzw = c < 4 ? fz0 : fz1;
    const uint32_t mw = c < 4 ? w1.z : w1.w;
    float tnx = fmaf((float)((nxw >> sh) & 0xffu), ax, bxn);
    float tfx = fmaf((float)((fxw >> sh) & 0xffu), ax, bxf);
    float tny = fmaf((float)((nyw >> sh) & 0xffu), ay, byn);
    float tfy = fmaf((float)((fyw >> sh) & 0xffu), ay, byf);
    float tnz = fmaf((float)((nzw >> sh) & 0xffu), az, bzn);
    float tfz = fmaf((float)((fzw >> sh) & 0xffu), az, bzf);
    float tmin = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, r.tnear));
    float tmax = fminf(fminf(fminf(tfx, tfy), tfz), tmax_ray);
    // an empty slot adds nothing even when its box is taken: its imask bit
    // and its device meta byte are 0
    if (tmin <= tmax) {
      // internal children (imask) set their octant bit; leaves OR in their
      // triangle bits (device meta: bits << 5 | offset; 0 for internal
      // children)
      hits |= ((imask >> c) & 1u) << (c ^ r.sstar);
      tmask |= ((mw >> (sh + 5)) & 7u) << ((mw >> sh) & 31u);
    }
  }
  s.gbase = w1.x;
  s.gbits = hits | (imask << 8);
  tri_base = w1.y;
  tmask_out = tmask;
  return false;
}

// Second half of a traversal step, one lane alone: test the node's
// triangles in ascending order. Returns true when ANY and a hit was taken.
// The loop has a single exit (an any-hit ray finishes the node's few
// triangles and then stops): with an early return from inside it, hipcc's
// gfx950 code for the any-hit variant carried the old group base instead of
// the child base out of nodes whose triangles all missed.
template <bool ANY, bool COUNT>
__device__ __forceinline__ bool trav_tris(const SceneView& sc, const TravRay& r, TravState& s,
                                          uint32_t tri_base, uint32_t tmask) {
  bool found = false;
  float rx, ry, rz;
  rot3(r.kz, r.ox, r.oy, r.oz, rx, ry, rz);
  while (tmask) {
    uint32_t j = __builtin_ctz(tmask);
    tmask &= tmask - 1u;
    const float4* tp = sc.tris + 3 * (size_t)(tri_base + j);
    float4 a = tp[0], b = tp[1], c = tp[2];
    if (COUNT) ++s.tris;
    float t, u, v;
    if (tri_test(r.kz, rx, ry, rz, r.Sx, r.Sy, r.Sz, a, b, c, t, u, v)) {
      int id = (int)as_uint(a.w);
      if (better_hit(t, id, r.tnear, s.tbest, s.best)) {
        s.tbest = t;
        s.best = id;
        s.bu = u;
        s.bv = v;
        found = true;
      }
    }
  }
  return ANY && found;
}

// One traversal step: visit one node and test its triangles. Returns true
// when the ray is finished (stack empty, or any-hit found).
template <bool ANY, bool COUNT, class Stack>
__device__ __forceinline__ bool trav_step(const SceneView& sc, const TravRay& r,
                                          TravState& s, Stack& stack, bool& overflow) {
  uint32_t tri_base, tmask;
  if (trav_node<COUNT>(sc, r, s, stack, overflow, tri_base, tmask)) return true;
  return trav_tris<ANY, COUNT>(sc, r, s, tri_base, tmask);
}

}  // namespace shb

#endif  // SHB_TRAVERSE_H_
