// ORACLE (test infrastructure only) — small vector math with the evaluation
// order of the Eigen expressions used by the reference (DaviesX/sh-baker).
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// the oracle, and only as the checker. Compiled with -ffp-contract=off so
// every a*b+c below is two roundings, as written; fused operations are
// spelled std::fma where the GPU kernels use fmaf by design.
//
// Eigen fixed-size reductions (dot, squaredNorm, maxCoeff, sum) unroll as
// e0 op (e1 op e2) (redux_novec_unroller halves the range); the helpers below
// keep that order so the oracle and the kernels round identically.
#ifndef SHB_ORACLE_MATH_H_
#define SHB_ORACLE_MATH_H_

#include <algorithm>
#include <cmath>
#include <cstdint>

namespace orc {

struct V2 {
  float x = 0.f, y = 0.f;
};
struct V3 {
  float x = 0.f, y = 0.f, z = 0.f;
};
struct V4 {
  float x = 0.f, y = 0.f, z = 0.f, w = 0.f;
};

inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3& operator+=(V3& a, V3 b) {
  a = a + b;
  return a;
}
inline V3 cmul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
inline float sqnorm(V3 a) { return a.x * a.x + (a.y * a.y + a.z * a.z); }
inline float norm(V3 a) { return std::sqrt(sqnorm(a)); }
// Eigen MatrixBase::normalized(): n / sqrt(|n|^2) when |n|^2 > 0.
inline V3 normalized(V3 a) {
  float z = sqnorm(a);
  if (z > 0.f) {
    float s = std::sqrt(z);
    return a / s;
  }
  return a;
}
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
          a.x * b.y - a.y * b.x};
}
inline float maxc(V3 a) { return std::max(a.x, std::max(a.y, a.z)); }
inline bool is_zero(V3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
inline float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

inline V2 operator+(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
inline V2 operator-(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
inline V2 operator*(V2 a, float s) { return {a.x * s, a.y * s}; }
inline V2 operator*(float s, V2 a) { return {s * a.x, s * a.y}; }
inline float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }

inline V4 operator+(V4 a, V4 b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w};
}
inline V4 operator*(V4 a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }

// -------------------------------------------------------------- RNG
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
// increments 0x9E3779B9 / 0xBB67AE85.
inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                          uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// xoshiro128** 1.1 (Blackman & Vigna): 32-bit output rotl(s1*5, 7)*9.
struct Xoshiro128ss {
  uint32_t s[4];
  static uint32_t rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
  uint32_t next() {
    const uint32_t result = rotl(s[1] * 5u, 7) * 9u;
    const uint32_t t = s[1] << 9;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 11);
    return result;
  }
  // Uniform float in [0, 1) with 24 bits.
  float uniform() { return (float)(next() >> 8) * 0x1p-24f; }
  // Two uniforms in [0, 1) with 16 bits each from one draw (high half, low
  // half): the GPU stream's barycentric pair of an area light.
  void uniform_pair16(float& a, float& b) {
    const uint32_t x = next();
    a = (float)(x >> 16) * 0x1p-16f;
    b = (float)(x & 0xffffu) * 0x1p-16f;
  }
  // The pair and, from the same step, the 24-bit uniform of xoshiro128+'s
  // output s[0] + s[3] (before the update): the GPU stream's barycentric
  // pair and pick uniform of an area light in the one-sample NEE.
  void uniform_pair16_and_pick(float& a, float& b, float& u) {
    const uint32_t plus = s[0] + s[3];
    uniform_pair16(a, b);
    u = (float)(plus >> 8) * 0x1p-24f;
  }
  // Uniform integer in [0, n) by 32x32 -> 64 multiply-shift.
  uint32_t index(uint32_t n) {
    return (uint32_t)(((uint64_t)next() * (uint64_t)n) >> 32);
  }
};

}  // namespace orc

#endif  // SHB_ORACLE_MATH_H_
