// Shared host/device definitions of the MI355X SH-lightmap baker: device
// scene layout, counter-based random stream, and the small vector math of
// the bake. All HIP code is compiled with -ffp-contract=off; fused ops are
// written as fmaf where the design wants them (traversal), so results are
// reproducible bit for bit by a host restatement.
#ifndef SHB_COMMON_H_
#define SHB_COMMON_H_

#include <stdint.h>

#ifdef SHB_HOST_ONLY
// Host-only build of the traversal / shading code (CPU checks of the exact
// device arithmetic): minimal stand-ins for the HIP vector types.
#include <cmath>
#include <cstring>
#define __host__
#define __device__
#define __forceinline__ inline
struct float2 { float x, y; };
struct float4 { float x, y, z, w; };
struct uint2 { uint32_t x, y; };
struct uint4 { uint32_t x, y, z, w; };
struct int4 { int32_t x, y, z, w; };
inline float4 make_float4(float x, float y, float z, float w) { return {x, y, z, w}; }
inline uint2 make_uint2(uint32_t x, uint32_t y) { return {x, y}; }
inline int4 make_int4(int x, int y, int z, int w) { return {x, y, z, w}; }
#include <math.h>
#include <algorithm>
using std::max;
using std::min;
using std::signbit;
inline uint32_t __builtin_popcount_u(uint32_t x) { return (uint32_t)__builtin_popcount(x); }
#else
#include <hip/hip_runtime.h>
#endif

#define SHB_HD __host__ __device__ __forceinline__

namespace shb {

// ------------------------------------------------------------ limits
constexpr int kMaxBounces = 15;       // depth fits 4 bits of the RNG counter
constexpr int kMaxSamples = 65535;    // sample fits 16 bits of the counter
constexpr int kMaxLightSamples = 16;  // light samples per NEE (a pick bit mask)
constexpr int kStackLds = 16;         // traversal stack entries (LDS)
// nodes of one traversed tree (a scene's, or the joined trees of a batch):
// node addresses are formed with 24-bit multiplies; above 2^24 triangle
// records the wave-spread triangle tests are off (same reason)
constexpr uint32_t kMaxTreeNodes = 1u << 24;
constexpr uint32_t kWaveTrisMaxRecords = 1u << 24;
constexpr uint32_t kNoLeaf = 0xffffffffu;  // AreaTri::leaf of a triangle outside the BVH
constexpr float kRayEps = 0.001f;     // baker.cpp:118, 122, light.h:40,83

// ------------------------------------------------------------ BVH node
// Compressed 8-wide node (after Ylitie, Karras, Laine, "Efficient Incoherent
// Ray Traversal on GPUs Through Compressed Wide BVHs", HPG 2017): child boxes
// quantised to 8 bits per plane on a per-node power-of-two grid anchored at
// (px,py,pz). 80 bytes = five 16-byte loads.
//   meta[s] == 0            : empty slot
//   meta[s] & 0x80          : internal child, rank (meta & 7) among the
//                             node's internal children -> child_base + rank
//   otherwise               : leaf, count = meta >> 5 (1..3) triangles at
//                             tri_base + (meta & 31)
// Slots are arranged so that visiting slot s in order of s ^ sstar, where
// sstar is the slot opposite to the ray's direction octant, is roughly front
// to back.
struct alignas(16) CWNode {
  float px, py, pz;
  int8_t ex, ey, ez;
  uint8_t imask;
  uint32_t child_base;
  uint32_t tri_base;
  uint8_t meta[8];
  uint8_t qlox[8], qloy[8], qloz[8];
  uint8_t qhix[8], qhiy[8], qhiz[8];
};
static_assert(sizeof(CWNode) == 80, "CWNode must be 80 bytes");

// The device form of a node: the same layout, but a leaf slot's meta byte
// holds its triangle bits instead of its count -- (2^count - 1) << 5 |
// offset, so the traversal ORs (meta >> 5) << offset into its triangle mask
// directly -- and an internal slot's meta byte is 0 (imask names internal
// slots and ranks them). The host tree and its export keep the encoding
// above.
inline CWNode device_node(const CWNode& n) {
  CWNode d = n;
  for (int s = 0; s < 8; ++s) {
    const uint32_t m = n.meta[s];
    if (m & 0x80u)
      d.meta[s] = 0;
    else if (m != 0u)
      d.meta[s] = (uint8_t)((((1u << (m >> 5)) - 1u) << 5) | (m & 31u));
  }
  return d;
}

// Bytes between consecutive nodes of the device node array (the host tree
// and its export are packed at 80).
#ifndef SHB_NODE_BYTES
#define SHB_NODE_BYTES 80
#endif
constexpr uint32_t kNodeBytes = SHB_NODE_BYTES;
static_assert(kNodeBytes >= 80 && kNodeBytes % 16 == 0, "node stride");

// ------------------------------------------------------------ lights
enum : int32_t { kLightPoint = 0, kLightDirectional = 1, kLightSpot = 2, kLightArea = 3 };

struct alignas(16) LightDev {
  float ic[3];         // intensity * color (light.h:48, 93, 143)
  int32_t type;
  float pos[3];
  float cos_inner;
  float dir[3];
  float cos_outer;
  float emit[3];       // area lights with a constant emitter: emission
  int32_t emit_const;  // 1: emit[] is valid; 0: look up the texture per sample
  uint32_t first_tri;  // area lights: first record in area_tris
  uint32_t num_tris;   // area lights: triangle count of the geometry (0: off)
  int32_t material;    // area lights: emitting material
  // selection-weight helpers: the largest channel of ic (of emit for a
  // constant emitter), and 1 when every channel of it is equal (a grey
  // light: its vertex weight needs only the largest BRDF channel)
  float wmax;
  int32_t grey;
  // lights from this one on (in scene order) that are single-triangle
  // area lights with a constant grey emitter; 0 when this one is not. The
  // NEE loop takes such runs in a branch-free form.
  uint32_t simple_run;
  int32_t pad;
};

// Area-light triangle record (object-space vertices, light.cpp:45-49): v0
// and the edges e1 = v1 - v0, e2 = v2 - v0 (float differences), so a sample
// point is v0 + u1 e1 + u2 e2 with two fused multiply-adds per axis. rec is
// a copy of the emitting triangle's record in the BVH's triangle array (the
// traversal's vertices, bit for bit), set at upload; leaf is that record's
// index there, kNoLeaf when the triangle has none (rec is then unused).
struct alignas(16) AreaTri {
  float v0[3];
  float pdf;           // max(1e-6, 1/ntri * 1/area)  (light.cpp:75-76)
  float e1[3];
  float has_uv;
  float e2[3];
  uint32_t leaf;
  float uv[6];
  float pad2[2];
  float4 rec[3];
};

// ------------------------------------------------------------ materials
struct alignas(16) MaterialDev {
  uint32_t tex_w, tex_h, tex_c;
  float emission;      // emission_intensity
  uint64_t tex_off;    // byte offset into the texel pool (tex_c == 0: none)
  uint32_t pad[2];
};

// Per-triangle shading data, indexed by global triangle id.
//   s0 = (n0.xyz, uv0.x), s1 = (n1.xyz, uv0.y), s2 = (n2.xyz, uv1.x),
//   s3 = (uv1.y, uv2.x, uv2.y, flags)  flags: bit0 vertex normals,
//   bit1 texture uvs. Without vertex normals n0 holds the normalised world
//   geometric normal.
// TriInfo = (material, emitting light or -1, geometry, local triangle).

// ------------------------------------------------------------ scene view
struct SceneView {
  const CWNode* nodes;
  const float4* tris;      // leaf order: 3 float4 per triangle, v0.w = id
  const float4* shade;     // 4 float4 per global triangle
  const int4* tri_info;
  const MaterialDev* mats;
  const uint8_t* texels;
  const float* srgb_lut;   // 256 entries, SRGBToLinear
  const LightDev* lights;
  const AreaTri* area_tris;
  const float* light_scale;   // 3 per light
  const float* albedo_scale;  // 3 per material
  int32_t num_lights;
  int32_t num_tris;
  int32_t has_nodes;
  int32_t emitter_test;    // 1: shadow queries ending on their emitter are
                           // answered by its triangle test (emitter_occludes)
};

// ------------------------------------------------------------ RNG
// Philox4x32-10 (Salmon et al., SC'11) seeds one xoshiro128** state per path
// vertex ("site"); counter = (texel, sample | depth<<16 | branch<<20,
// kind | (branch>>12)<<8, job), key = 64-bit seed. The branch id names the
// vertex's place in the tree of alpha splits (root 1; transmitted 2b + 1,
// reflected 2b), below 2^16 for 15 bounces.
enum : uint32_t { kSiteTexel = 0, kSiteVertex = 1 };

SHB_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

struct Rng {
  uint32_t s0, s1, s2, s3;
  SHB_HD static uint32_t rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
  SHB_HD uint32_t next() {
    const uint32_t result = rotl(s1 * 5u, 7) * 9u;
    const uint32_t t = s1 << 9;
    s2 ^= s0;
    s3 ^= s1;
    s1 ^= s2;
    s0 ^= s3;
    s2 ^= t;
    s3 = rotl(s3, 11);
    return result;
  }
  SHB_HD float uniform() { return (float)(next() >> 8) * 0x1p-24f; }
  // Two uniforms in [0, 1) with 16 bits each from one draw (high half,
  // low half): an area light's barycentric pair.
  SHB_HD void uniform_pair16(float& a, float& b) {
    const uint32_t x = next();
    a = (float)(x >> 16) * 0x1p-16f;
    b = (float)(x & 0xffffu) * 0x1p-16f;
  }
  // The pair above and, from the same step, a 24-bit uniform from the
  // state's other scrambler (xoshiro128+: s0 + s3, before the update): an
  // area light's barycentric pair and pick uniform for one state update.
  SHB_HD void uniform_pair16_and_pick(float& a, float& b, float& u) {
    const uint32_t plus = s0 + s3;
    uniform_pair16(a, b);
    u = (float)(plus >> 8) * 0x1p-24f;
  }
  SHB_HD uint32_t index(uint32_t n) {
    return (uint32_t)(((uint64_t)next() * (uint64_t)n) >> 32);
  }
};

SHB_HD Rng site_rng(uint64_t seed, uint32_t texel, uint32_t sample,
                    uint32_t depth, uint32_t branch, uint32_t kind,
                    uint32_t job) {
  uint32_t c[4] = {texel, sample | (depth << 16) | (branch << 20), kind | ((branch >> 12) << 8),
                   job};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  Rng r{c[0], c[1], c[2], c[3]};
  if ((c[0] | c[1] | c[2] | c[3]) == 0u) r.s0 = 1u;
  return r;
}

// A value the caller knows to be equal on every active lane, made visibly
// uniform to the compiler (scalar registers, scalar-cache loads).
//
// uniform_load(p, i): p[i] for a wave-uniform i from memory no kernel writes
// while it runs (scene tables), read through the constant address space so
// it can be a scalar load.
// const_load(p, i): p[i] from such memory where p or i may differ between
// lanes (a wave holding texels of several jobs): the compiler makes it a
// scalar load only where it proves both uniform, a vector load elsewhere.
#if defined(SHB_HOST_ONLY)
SHB_HD uint32_t uniform_index(uint32_t v) { return v; }
SHB_HD int uniform_index(int v) { return v; }
template <class T>
SHB_HD T uniform_load(const T* p, uint32_t i) { return p[i]; }
template <class T>
SHB_HD T const_load(const T* p, uint32_t i) { return p[i]; }
#else
template <class T>
__device__ __forceinline__ T uniform_load(const T* p, uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef const __attribute__((address_space(4))) T* cptr;
  return ((cptr)p)[__builtin_amdgcn_readfirstlane((int)i)];
#else
  return p[i];
#endif
}
template <class T>
__device__ __forceinline__ T const_load(const T* p, uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef const __attribute__((address_space(4))) T* cptr;
  return ((cptr)p)[i];
#else
  return p[i];
#endif
}
__device__ __forceinline__ uint32_t uniform_index(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ int uniform_index(int v) { return __builtin_amdgcn_readfirstlane(v); }
#endif

// ------------------------------------------------------------ math
// Reductions keep Eigen's fixed-size order e0 + (e1 + e2).
struct f3 {
  float x, y, z;
};
SHB_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
SHB_HD f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
SHB_HD f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
SHB_HD f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
SHB_HD f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
SHB_HD f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
SHB_HD f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
SHB_HD f3 cmul(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
SHB_HD float dot(f3 a, f3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
SHB_HD float sqnorm(f3 a) { return a.x * a.x + (a.y * a.y + a.z * a.z); }
SHB_HD f3 normalized(f3 a) {
  float z = sqnorm(a);
  if (z > 0.f) {
    float s = sqrtf(z);
    return a / s;
  }
  return a;
}
SHB_HD f3 cross(f3 a, f3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
SHB_HD float maxc(f3 a) {
  float m = a.y < a.z ? a.z : a.y;
  return a.x < m ? m : a.x;
}
SHB_HD bool is_zero(f3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
SHB_HD f3 f3_of(float4 v) { return {v.x, v.y, v.z}; }

// Bit casts usable on host and device.
SHB_HD float as_float(uint32_t u) {
  union {
    uint32_t u;
    float f;
  } x;
  x.u = u;
  return x.f;
}
SHB_HD uint32_t as_uint(float f) {
  union {
    uint32_t u;
    float f;
  } x;
  x.f = f;
  return x.u;
}

// 1 / sqrt(d2) as fl(1 / fl(sqrt(d2))) and the distance fl(sqrt(d2)), both
// correctly rounded. For d2 in [2^-96, 2^126] (distances 2^-48 .. 2^63) the
// hardware estimates (within one ulp) are corrected exactly as the general
// IEEE sequences correct them, without those sequences' range scaling and
// special-value steps: the same bits as sqrtf and 1.0f / x there (checked for
// every float of the range on the GPU); other d2 take sqrtf and the division.
#if defined(SHB_HOST_ONLY)
SHB_HD void inv_dist(float d2, float& dist, float& inv) {
  dist = std::sqrt(d2);
  inv = 1.0f / dist;
}
#else
__device__ __forceinline__ float sqrt_core(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sm = as_float(as_uint(s) - 1u);
  const float sp = as_float(as_uint(s) + 1u);
  const float t = fmaf(-sm, s, x) <= 0.f ? sm : s;
  return fmaf(-sp, s, x) > 0.f ? sp : t;
}
__device__ __forceinline__ float rcp_core(float d) {
  const float y0 = __builtin_amdgcn_rcpf(d);
  const float y1 = fmaf(fmaf(-d, y0, 1.0f), y0, y0);
  const float q1 = fmaf(fmaf(-d, y1, 1.0f), y1, y1);
  return fmaf(fmaf(-d, q1, 1.0f), y1, q1);
}
constexpr float kCoreLo = 0x1p-96f, kCoreHi = 0x1p126f;
__device__ __forceinline__ void inv_dist(float d2, float& dist, float& inv) {
  if (d2 >= kCoreLo && d2 <= kCoreHi) {
    dist = sqrt_core(d2);
    inv = rcp_core(dist);
  } else {
    dist = sqrtf(d2);
    inv = 1.0f / dist;
  }
}
#endif

}  // namespace shb

#endif  // SHB_COMMON_H_
