// pm_scene.h -- scene description, ray/primitive intersection, materials,
// direction sampling and the participating medium.  Shared per-thread code
// (GPU kernels and host oracle).
//
// Reference counterparts (photonMappingKernel.cu):
//   scene globals            :59-73   (open box of 5 axis-aligned planes, spheres, point light)
//   colours per plane        :605-617 (plane 0 green, plane 2 red, others white)
//   checkDistance/raySphere/rayPlane/rayObject/raytrace  :106-165, :223-241
//   sphereNormal/planeNormal/surfaceNormal               :173-209
//   reflect3 / refract3                                  :620-668
//   medium handling in emission                          :1239-1272
// The new scene is the unit room [0,1]^3 with the same topology (5 walls, +x
// green, -x red, open front), diffuse / mirror / dielectric spheres, and a
// physically based homogeneous or heterogeneous medium.
#pragma once

#include "pm_core.h"

namespace pm {

enum Material : int { MAT_DIFFUSE = 0, MAT_MIRROR = 1, MAT_DIELECTRIC = 2 };
enum MediumType : int { MED_NONE = 0, MED_HOMOGENEOUS = 1, MED_NOISE = 2 };

struct Sphere {
  float cx, cy, cz, r;
  float ar, ag, ab;  // albedo (diffuse) / tint (mirror, dielectric)
  int mat;
  float ior;
  float pad0, pad1, pad2;
};

struct Medium {
  int type;
  float sigma_s, sigma_a;  // scale of the density field (homogeneous: the value)
  float g;                 // Henyey-Greenstein asymmetry (0: isotropic)
  int res;                 // density grid vertices per axis
  float majorant;          // upper bound of sigma_t over the box
  const float* density;    // res^3 vertex values in [0,1], x fastest
};

// The 5 walls of the unit room, faces in the reference's plane order
// (photonMappingKernel.cu:67-73): 0: x=1 (green), 1: y=0 (floor),
// 2: x=0 (red), 3: y=1 (ceiling), 4: z=1 (back).  The z=0 face is open.
struct Scene {
  float wall_albedo[5][3];
  int nspheres;
  const Sphere* spheres;
  float light[3];
  float light_power[3];
  Medium med;
};

PM_HD int face_axis(int f) { return f == 0 || f == 2 ? 0 : (f == 1 || f == 3 ? 1 : 2); }
PM_HD float face_value(int f) { return (f == 0 || f == 3 || f == 4) ? 1.0f : 0.0f; }
// Inward-facing normal sign along the face axis.
PM_HD float face_nsign(int f) { return (f == 0 || f == 3 || f == 4) ? -1.0f : 1.0f; }

enum HitKind : int { HIT_NONE = 0, HIT_WALL = 1, HIT_SPHERE = 2 };

template <class T>
struct Hit {
  T t;
  int kind;
  int id;
};

// Closest positive hit against the 5 wall rectangles and the spheres.
template <class T>
PM_HD Hit<T> intersect(const Scene& sc, V3<T> o, V3<T> d) {
  Hit<T> h;
  h.t = T(3.0e38);
  h.kind = HIT_NONE;
  h.id = -1;
  const T one = T(1), zero = T(0);
  for (int f = 0; f < 5; ++f) {
    int a = face_axis(f);
    T da = comp(d, a);
    if (da == zero) continue;
    T t = (T(face_value(f)) - comp(o, a)) / da;
    if (!(t > zero) || !(t < h.t)) continue;
    int b = a == 0 ? 1 : 0;
    int c = a == 2 ? 1 : 2;
    T pb = comp(o, b) + t * comp(d, b);
    T pc = comp(o, c) + t * comp(d, c);
    if (pb < zero || pb > one || pc < zero || pc > one) continue;
    h.t = t;
    h.kind = HIT_WALL;
    h.id = f;
  }
  for (int i = 0; i < sc.nspheres; ++i) {
    const Sphere& s = sc.spheres[i];
    V3<T> oc = o - v3<T>(T(s.cx), T(s.cy), T(s.cz));
    T b = dot(oc, d);
    T c2 = dot(oc, oc) - T(s.r) * T(s.r);
    T disc = b * b - c2;
    if (disc < zero) continue;
    T sq = sqrt_(disc);
    T t = -b - sq;
    if (!(t > zero)) t = -b + sq;
    if (!(t > zero) || !(t < h.t)) continue;
    h.t = t;
    h.kind = HIT_SPHERE;
    h.id = i;
  }
  return h;
}

// Geometric normal at a hit, pointing outward for spheres and into the room for walls.
template <class T>
PM_HD V3<T> hit_normal(const Scene& sc, const Hit<T>& h, V3<T> p) {
  if (h.kind == HIT_WALL) {
    int a = face_axis(h.id);
    T s = T(face_nsign(h.id));
    return v3<T>(a == 0 ? s : T(0), a == 1 ? s : T(0), a == 2 ? s : T(0));
  }
  const Sphere& s = sc.spheres[h.id];
  V3<T> n = p - v3<T>(T(s.cx), T(s.cy), T(s.cz));
  return n * (T(1) / T(s.r));
}

PM_HD int hit_material(const Scene& sc, int kind, int id) {
  return kind == HIT_WALL ? (int)MAT_DIFFUSE : sc.spheres[id].mat;
}

template <class T>
PM_HD V3<T> hit_albedo(const Scene& sc, int kind, int id) {
  if (kind == HIT_WALL) return v3<T>(T(sc.wall_albedo[id][0]), T(sc.wall_albedo[id][1]), T(sc.wall_albedo[id][2]));
  const Sphere& s = sc.spheres[id];
  return v3<T>(T(s.ar), T(s.ag), T(s.ab));
}

// Distance along d from o (inside the room) to the open z=0 face; +inf-like if not heading out.
template <class T>
PM_HD T t_exit_front(V3<T> o, V3<T> d) {
  if (d.z < T(0)) return -o.z / d.z;
  return T(0);
}

// ---------------------------------------------------------------- directions
template <class T>
PM_HD V3<T> reflect(V3<T> d, V3<T> n) {
  T k = T(2) * dot(d, n);
  return d - n * k;
}

// Orthonormal basis around unit n (Duff et al., "Building an Orthonormal
// Basis, Revisited", JCGT 2017).
template <class T>
PM_HD void onb(V3<T> n, V3<T>* t, V3<T>* b) {
  T sign = n.z >= T(0) ? T(1) : T(-1);
  T a = T(-1) / (sign + n.z);
  T bb = n.x * n.y * a;
  *t = v3<T>(T(1) + sign * n.x * n.x * a, sign * bb, -sign * n.x);
  *b = v3<T>(bb, sign + n.y * n.y * a, -n.y);
}

PM_HD void sincos_t(float u, float* s, float* c) { sincos_turns(u, s, c); }
#if !defined(__HIP_DEVICE_COMPILE__)
inline void sincos_t(double u, double* s, double* c) {
  const double tp = 6.283185307179586476925286766559;
  *s = ::sin(tp * u);
  *c = ::cos(tp * u);
}
#endif

// Uniform direction on the unit sphere.
template <class T>
PM_HD V3<T> sample_sphere(T u0, T u1) {
  T z = T(1) - T(2) * u0;
  T r = sqrt_(fmax_(T(0), T(1) - z * z));
  T s, c;
  sincos_t(u1, &s, &c);
  return v3<T>(r * c, r * s, z);
}

// Cosine-weighted direction in the hemisphere around unit n.
template <class T>
PM_HD V3<T> sample_cosine(V3<T> n, T u0, T u1) {
  T r = sqrt_(u0);
  T s, c;
  sincos_t(u1, &s, &c);
  T z = sqrt_(fmax_(T(0), T(1) - u0));
  V3<T> t, b;
  onb(n, &t, &b);
  return (t * (r * c) + b * (r * s)) + n * z;
}

// Henyey-Greenstein: cos(theta) between incoming travel direction d and the
// scattered travel direction; g = 0 is isotropic.
template <class T>
PM_HD T hg_cos(T g, T u) {
  T ag = g < T(0) ? -g : g;
  if (ag < T(1e-3)) return T(1) - T(2) * u;
  T sq = (T(1) - g * g) / (T(1) - g + T(2) * g * u);
  T c = (T(1) + g * g - sq * sq) / (T(2) * g);
  return fmin_(T(1), fmax_(T(-1), c));
}

template <class T>
PM_HD V3<T> sample_hg(V3<T> d, T g, T u0, T u1) {
  T ct = hg_cos(g, u0);
  T st = sqrt_(fmax_(T(0), T(1) - ct * ct));
  T s, c;
  sincos_t(u1, &s, &c);
  V3<T> t, b;
  onb(d, &t, &b);
  return (t * (st * c) + b * (st * s)) + d * ct;
}

// HG phase function value, cos_theta between incoming and outgoing travel directions.
template <class T>
PM_HD T phase_hg(T cos_theta, T g) {
  const T inv4pi = T(0.0795774715459476678844418816863);
  T den = T(1) + g * g - T(2) * g * cos_theta;
  return inv4pi * (T(1) - g * g) / (den * sqrt_(den));
}
// FP32: den^(-3/2) from y = den^(-1/2) by three Newton steps
// y <- y (3/2 - den/2 y^2), seeded by halving the exponent in the integer
// bits (C. Lomont, "Fast Inverse Square Root", 2003, constant 0x5f375a86:
// relative error below 1.8e-3, then 4.7e-6, then the rounding floor).  The
// result is within 6e-7 of den^(-3/2) for every den in [2^-48, 4] (|g| < 1),
// against 1.5e-7 for the correctly rounded sqrt and division it replaces, at
// about half their instructions on the GPU; plain multiplies and subtractions
// only, so the CPU and the GPU give the same bits.
PM_HD float phase_hg(float cos_theta, float g) {
  const float inv4pi = 0.0795774715459476678844418816863f;
  const float den = 1.0f + g * g - 2.0f * g * cos_theta;
  const float h = 0.5f * den;
  float y = u2f(0x5f375a86u - (f2u(den) >> 1));
  y = y * (1.5f - h * (y * y));
  y = y * (1.5f - h * (y * y));
  y = y * (1.5f - h * (y * y));
  return inv4pi * (1.0f - g * g) * ((y * y) * y);
}

// Unpolarised Fresnel reflectance for a dielectric interface; cos_i > 0 is the
// cosine on the incident side, eta = n_incident / n_transmitted.  Returns 1 on
// total internal reflection and sets *cos_t otherwise.
template <class T>
PM_HD T fresnel(T cos_i, T eta, T* cos_t) {
  T sin2t = eta * eta * (T(1) - cos_i * cos_i);
  if (sin2t >= T(1)) {
    *cos_t = T(0);
    return T(1);
  }
  T ct = sqrt_(T(1) - sin2t);
  *cos_t = ct;
  T rs = (eta * cos_i - ct) / (eta * cos_i + ct);
  T rp = (cos_i - eta * ct) / (cos_i + eta * ct);
  return (rs * rs + rp * rp) * T(0.5);
}

// Refracted direction for unit d hitting a surface with unit normal n facing
// the incident side (dot(d, n) < 0).
template <class T>
PM_HD V3<T> refract(V3<T> d, V3<T> n, T eta, T cos_i, T cos_t) {
  return d * eta + n * (eta * cos_i - cos_t);
}

// ---------------------------------------------------------------- medium
// Trilinear interpolation of the vertex grid; p is clamped to the room.
template <class T>
PM_HD T density_at(const Medium& m, V3<T> p) {
  if (m.type != MED_NOISE) return T(1);
  const int R = m.res;
  const T s = T(R - 1);
  T fx = fmin_(fmax_(p.x, T(0)), T(1)) * s;
  T fy = fmin_(fmax_(p.y, T(0)), T(1)) * s;
  T fz = fmin_(fmax_(p.z, T(0)), T(1)) * s;
  int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  if (ix > R - 2) ix = R - 2;
  if (iy > R - 2) iy = R - 2;
  if (iz > R - 2) iz = R - 2;
  T tx = fx - T(ix), ty = fy - T(iy), tz = fz - T(iz);
  const float* g = m.density + ((size_t)iz * R + iy) * R + ix;
  const size_t sy = (size_t)R, sz = (size_t)R * R;
  T c00 = T(g[0]) + tx * (T(g[1]) - T(g[0]));
  T c10 = T(g[sy]) + tx * (T(g[sy + 1]) - T(g[sy]));
  T c01 = T(g[sz]) + tx * (T(g[sz + 1]) - T(g[sz]));
  T c11 = T(g[sz + sy]) + tx * (T(g[sz + sy + 1]) - T(g[sz + sy]));
  T c0 = c00 + ty * (c10 - c00);
  T c1 = c01 + ty * (c11 - c01);
  return c0 + tz * (c1 - c0);
}

template <class T> PM_HD T sigma_t_at(const Medium& m, V3<T> p) {
  return (T(m.sigma_s) + T(m.sigma_a)) * density_at(m, p);
}

// Value-noise vertex value of the density grid at vertex (i, j, k): trilinear
// interpolation of a (freq+1)^3 lattice of Philox uniforms (seed, TAG_NOISE).
PM_HD float noise_vertex(uint32_t seed, int res, int freq, int i, int j, int k) {
  const float s = (float)freq / (float)(res - 1);
  float fx = (float)i * s, fy = (float)j * s, fz = (float)k * s;
  int lx = (int)fx, ly = (int)fy, lz = (int)fz;
  if (lx > freq - 1) lx = freq - 1;
  if (ly > freq - 1) ly = freq - 1;
  if (lz > freq - 1) lz = freq - 1;
  float tx = fx - (float)lx, ty = fy - (float)ly, tz = fz - (float)lz;
  const uint32_t L = (uint32_t)freq + 1u;
  float v[8];
  for (int c = 0; c < 8; ++c) {
    uint32_t xx = (uint32_t)(lx + (c & 1)), yy = (uint32_t)(ly + ((c >> 1) & 1)), zz = (uint32_t)(lz + (c >> 2));
    v[c] = u01(rng(seed, (zz * L + yy) * L + xx, 0u, 0u, TAG_NOISE).x);
  }
  float c00 = v[0] + tx * (v[1] - v[0]);
  float c10 = v[2] + tx * (v[3] - v[2]);
  float c01 = v[4] + tx * (v[5] - v[4]);
  float c11 = v[6] + tx * (v[7] - v[6]);
  float c0 = c00 + ty * (c10 - c00);
  float c1 = c01 + ty * (c11 - c01);
  return c0 + tz * (c1 - c0);
}

}  // namespace pm
