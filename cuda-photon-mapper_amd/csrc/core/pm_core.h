// pm_core.h -- per-thread numeric core shared by the gfx950 kernels and the
// host-compiled oracle.
//
// Everything here is plain scalar code that compiles both as HIP device code
// (hipcc, gfx950) and as host C++ (g++).  Both sides are built with
// -ffp-contract=off and without fast-math, and every function below uses only
// IEEE add/sub/mul/div/sqrt and integer/bit operations, so a given input gives
// the same bits on the GPU and on the host cores.  Transcendentals are the
// project's own (range reduction + truncated series written out here), never a
// vendor math library.
//
// Replaces the reference's cutil_math.h vector helpers (included at
// photonMappingKernel.cu:4) and its global-state multiply-with-carry generator
// (photonMappingKernel.cu:1026-1056) with a stateless counter-based generator.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PM_HD __host__ __device__ inline
#else
#include <math.h>
#define PM_HD inline
#endif

namespace pm {

// ---------------------------------------------------------------- bit casts
PM_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
PM_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

// Order-preserving map float -> uint32 (total order, -0 < +0; no NaNs expected).
PM_HD uint32_t float_key(float f) {
  uint32_t u = f2u(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

PM_HD float fmin_(float a, float b) { return b < a ? b : a; }
PM_HD float fmax_(float a, float b) { return b > a ? b : a; }
PM_HD double fmin_(double a, double b) { return b < a ? b : a; }
PM_HD double fmax_(double a, double b) { return b > a ? b : a; }

// IEEE square root (correctly rounded on both sides: sqrtss on x86, the
// correctly-rounded expansion hipcc emits for gfx950 by default).
PM_HD float sqrt_(float x) { return __builtin_sqrtf(x); }
PM_HD double sqrt_(double x) { return __builtin_sqrt(x); }
// Round to nearest, ties to even (roundss / v_rndne_f32).
PM_HD float rint_(float x) { return __builtin_rintf(x); }
PM_HD float floor_(float x) { return __builtin_floorf(x); }
PM_HD double floor_(double x) { return __builtin_floor(x); }

// ---------------------------------------------------------------- vectors
template <class T>
struct V3 {
  T x, y, z;
};
typedef V3<float> V3f;
typedef V3<double> V3d;

template <class T> PM_HD V3<T> v3(T x, T y, T z) { V3<T> r; r.x = x; r.y = y; r.z = z; return r; }
template <class T> PM_HD V3<T> operator+(V3<T> a, V3<T> b) { return v3<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> PM_HD V3<T> operator-(V3<T> a, V3<T> b) { return v3<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> PM_HD V3<T> operator-(V3<T> a) { return v3<T>(-a.x, -a.y, -a.z); }
template <class T> PM_HD V3<T> operator*(V3<T> a, T s) { return v3<T>(a.x * s, a.y * s, a.z * s); }
template <class T> PM_HD V3<T> operator*(T s, V3<T> a) { return v3<T>(a.x * s, a.y * s, a.z * s); }
template <class T> PM_HD V3<T> mul(V3<T> a, V3<T> b) { return v3<T>(a.x * b.x, a.y * b.y, a.z * b.z); }
// Fixed evaluation order, no fused multiply-add: ((x*x' + y*y') + z*z').
template <class T> PM_HD T dot(V3<T> a, V3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T> PM_HD V3<T> cross(V3<T> a, V3<T> b) {
  return v3<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// Normalise with a true square root and one correctly-rounded reciprocal (no rsqrt).
template <class T> PM_HD V3<T> normalize(V3<T> a) {
  T inv = T(1) / sqrt_(dot(a, a));
  return a * inv;
}
template <class T> PM_HD T comp(V3<T> a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
template <class U, class T> PM_HD V3<U> cast3(V3<T> a) { return v3<U>((U)a.x, (U)a.y, (U)a.z); }

// ---------------------------------------------------------------- own transcendentals
// sin(2*pi*u) and cos(2*pi*u) ("turns").  Reduction to the nearest quarter turn
// is exact for |u| < 2^21 (4u and the subtraction are exact in FP32); the
// reduced argument f lies in [-1/8, 1/8] turn, i.e. |theta| <= pi/4, where the
// Taylor series of sin to theta^9 and of cos to theta^10 are below 2e-9
// relative truncation error.  Coefficients are (2*pi)^n/n! rounded to FP32.
#define PM_S1 6.283185307179586f
#define PM_S3 (-41.341702240399755f)
#define PM_S5 81.60524927607504f
#define PM_S7 (-76.70585975306136f)
#define PM_S9 42.058693944897634f
#define PM_C2 (-19.739208802178716f)
#define PM_C4 64.93939402266828f
#define PM_C6 (-85.45681720669371f)
#define PM_C8 60.24464137187664f
#define PM_C10 (-26.426256783374388f)

PM_HD void sincos_turns(float u, float* s_out, float* c_out) {
  float q = rint_(u * 4.0f);
  float f = u - q * 0.25f;
  int iq = ((int)q) & 3;
  float f2 = f * f;
  float sp = f * (PM_S1 + f2 * (PM_S3 + f2 * (PM_S5 + f2 * (PM_S7 + f2 * PM_S9))));
  float cp = 1.0f + f2 * (PM_C2 + f2 * (PM_C4 + f2 * (PM_C6 + f2 * (PM_C8 + f2 * PM_C10))));
  // quadrant iq: (s, c) = (sp, cp), (cp, -sp), (-sp, -cp), (-cp, sp) for iq =
  // 0..3, by selects and sign-bit flips (no branches, the same bits)
  const bool odd = (iq & 1) != 0;
  const float s = odd ? cp : sp, c = odd ? sp : cp;
  *s_out = u2f(f2u(s) ^ ((uint32_t)(iq & 2) << 30));
  *c_out = u2f(f2u(c) ^ ((uint32_t)((iq + 1) & 2) << 30));
}

// Natural log for finite x > 0.  x = m * 2^e with m in [sqrt(1/2), sqrt(2));
// log m = 2*atanh(s), s = (m-1)/(m+1), |s| <= 0.1716, series to s^9
// (truncation < 1e-9 relative); e*ln2 split in a 15-bit head (exact products
// for |e| < 512) and a tail.
#define PM_LN2_HI 0.693145751953125f
#define PM_LN2_LO 1.4286068202862268e-06f
PM_HD float log_(float x) {
  uint32_t b = f2u(x);
  int e = (int)((b >> 23) & 0xffu);
  if (e == 0) {  // subnormal input: rescale by 2^24
    x = x * 16777216.0f;
    b = f2u(x);
    e = (int)((b >> 23) & 0xffu) - 24;
  }
  e -= 127;
  float m = u2f((b & 0x007fffffu) | 0x3f800000u);
  if (m > 1.41421356237f) {
    m = m * 0.5f;
    e += 1;
  }
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  float p = s2 * (0.666666666666666667f + s2 * (0.4f + s2 * (0.285714285714285714f + s2 * 0.222222222222222222f)));
  float lm = (s + s) + s * p;
  float fe = (float)e;
  return fe * PM_LN2_HI + (fe * PM_LN2_LO + lm);
}

// exp(x) in FP32: k = rint(x/ln2), r = x - k*ln2 (Cody-Waite, two-part ln2),
// |r| <= 0.347, Taylor to r^7 (truncation < 6e-9 relative), then *2^k through
// the exponent field (two steps below 2^-126).
PM_HD float exp_(float x) {
  if (x < -103.0f) return 0.0f;
  if (x > 88.72f) return u2f(0x7f800000u);
  float kf = rint_(x * 1.44269504088896341f);
  float r = (x - kf * PM_LN2_HI) - kf * PM_LN2_LO;
  float p = 1.0f + r * (1.0f + r * (0.5f + r * (0.166666666666666667f + r * (0.0416666666666666667f +
                    r * (0.00833333333333333333f + r * (0.00138888888888888889f + r * 0.000198412698412698413f))))));
  int k = (int)kf;
  if (k >= -125) return p * u2f((uint32_t)(k + 127) << 23);
  // k in [-149, -126]: split the scaling so every intermediate stays normal.
  return (p * u2f((uint32_t)(k + 127 + 64) << 23)) * u2f((uint32_t)(127 - 64) << 23);
}

// ---------------------------------------------------------------- Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3"
// (SC'11).  Stateless: (counter, key) -> 4 x uint32.  Counter layout used by
// this project: (entity id, event, sub-index, purpose tag); key = (seed, 0).
struct U4 {
  uint32_t x, y, z, w;
};
PM_HD U4 u4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { U4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }

PM_HD U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  for (int i = 0; i < 10; ++i) {
    if (i) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c.x;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c.z;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = u4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
  }
  return c;
}

// Top 24 bits -> float in [0, 1 - 2^-24]; exact.
PM_HD float u01(uint32_t u) { return (float)(u >> 8) * 5.9604644775390625e-08f; }

// Purpose tags (counter word 3).
enum : uint32_t {
  TAG_EMIT = 1u,     // (photon id, 0, 0): emission direction
  TAG_FLIGHT = 2u,   // (photon id, event, iteration): free-flight / delta tracking
  TAG_SCATTER = 3u,  // (photon id, event, 0): new direction, roulette, Fresnel choice
  TAG_CAMERA = 16u,  // (pixel id, sample, 0): pixel jitter, march offset
  TAG_EYE = 17u,     // (pixel id, sample, depth): Fresnel choice on the eye path
  TAG_NOISE = 32u,   // (lattice id, 0, 0): value-noise lattice
  TAG_SCENE = 48u,   // (scene id, object, 0): random scene generation
  TAG_CACHE = 64u,   // (frame, sample id, 0): cache training sample selection
};

PM_HD U4 rng(uint32_t seed, uint32_t id, uint32_t ev, uint32_t sub, uint32_t tag) {
  return philox4x32_10(u4(id, ev, sub, tag), seed, 0u);
}

// ---------------------------------------------------------------- compensated sum
// Neumaier's variant of Kahan summation (FP32 radiance accumulation).
template <class T>
struct NSum {
  T s, c;
};
template <class T> PM_HD void nsum_init(NSum<T>& a) { a.s = T(0); a.c = T(0); }
template <class T> PM_HD void nsum_add(NSum<T>& a, T v) {
  T t = a.s + v;
  T as = a.s < T(0) ? -a.s : a.s;
  T av = v < T(0) ? -v : v;
  if (as >= av) a.c = a.c + ((a.s - t) + v);
  else a.c = a.c + ((v - t) + a.s);
  a.s = t;
}
template <class T> PM_HD T nsum_get(const NSum<T>& a) { return a.s + a.c; }

template <class T>
struct NSum3 {
  NSum<T> x, y, z;
};
template <class T> PM_HD void nsum3_init(NSum3<T>& a) { nsum_init(a.x); nsum_init(a.y); nsum_init(a.z); }
template <class T> PM_HD void nsum3_add(NSum3<T>& a, V3<T> v) { nsum_add(a.x, v.x); nsum_add(a.y, v.y); nsum_add(a.z, v.z); }
template <class T> PM_HD V3<T> nsum3_get(const NSum3<T>& a) { return v3<T>(nsum_get(a.x), nsum_get(a.y), nsum_get(a.z)); }

// ---------------------------------------------------------------- bf16 storage
// Round-to-nearest-even FP32 -> BF16 (finite inputs) and back.
PM_HD uint16_t f32_to_bf16(float f) {
  uint32_t u = f2u(f);
  uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
PM_HD float bf16_to_f32(uint16_t h) { return u2f((uint32_t)h << 16); }

}  // namespace pm
