// pm_transport.h -- per-photon trace and per-eye-sample path, shared by the
// gfx950 kernels and the host oracle ("the same per-thread logic on the same
// RNG stream").
//
// Reference counterparts (photonMappingKernel.cu):
//   emitPhotons (emission, medium steps, caustic aiming, bounce loop) :1215-1375
//   storePhoton / storeVolumePhoton / shadowPhoton                    :1147-1196
//   computePixelColor (eye ray, ray-march, specular chains, gather)   :926-1017
//   handleReflection / handleRefraction chains                         :673-827
// The physics here follows the new method (Jensen-style photon records,
// Russian roulette, free-flight / delta tracking, HG phase) instead of the
// reference's fixed-weight voxel splats; the pipeline shape (emit -> store ->
// bounce; eye ray -> march -> specular chain -> gather) is the reference's.
#pragma once

#include "pm_scene.h"

namespace pm {

struct TraceParams {
  uint32_t seed;
  int max_bounces;     // scattering events per photon (8)
  float rr_survive;    // Russian-roulette survival probability (0.8)
  float eps;           // ray-origin offset along the normal
  int max_delta_iters; // bound on delta-tracking iterations per segment
  int media_on;
};

// Per-photon flux bookkeeping, in units of the photon's initial power (1 per
// channel): absorbed[3], escaped[3], rr_gain[3].  For every photon,
// 1 + rr_gain == absorbed + escaped holds term by term (each update adds the
// exact difference of two FP32 values), so the totals close up to FP64 rounding.
struct Tally9 {
  double v[9];
};
PM_HD void tally_zero(Tally9& t) {
  for (int i = 0; i < 9; ++i) t.v[i] = 0.0;
}
PM_HD void tally_add3(Tally9& t, int base, V3f a, V3f b) {  // += a - b (exact per term)
  t.v[base + 0] += (double)a.x - (double)b.x;
  t.v[base + 1] += (double)a.y - (double)b.y;
  t.v[base + 2] += (double)a.z - (double)b.z;
}

// Free-flight sampling in the medium along [0, tmax).  Homogeneous: exponential
// distance.  Density field: delta (Woodcock) tracking against the majorant.
PM_HD bool sample_collision(const Medium& m, uint32_t seed, int max_iters, uint32_t pid, int ev,
                            V3f o, V3f d, float tmax, float* tc) {
  if (m.type == MED_HOMOGENEOUS) {
    U4 r = rng(seed, pid, (uint32_t)ev, 0u, TAG_FLIGHT);
    float st = m.sigma_s + m.sigma_a;
    float t = -log_(1.0f - u01(r.x)) / st;
    if (t < tmax) {
      *tc = t;
      return true;
    }
    return false;
  }
  float t = 0.0f;
  for (int it = 0; it < max_iters; ++it) {
    U4 r = rng(seed, pid, (uint32_t)ev, (uint32_t)it, TAG_FLIGHT);
    t = t - log_(1.0f - u01(r.x)) / m.majorant;
    if (!(t < tmax)) return false;
    V3f x = o + d * t;
    if (u01(r.y) * m.majorant < sigma_t_at(m, x)) {
      *tc = t;
      return true;
    }
  }
  return false;
}

// Roulette after an interaction: returns false if the photon is absorbed.
PM_HD bool roulette(float u, float p, V3f& beta, Tally9& tal) {
  if (u >= p) {
    tally_add3(tal, 0, beta, v3(0.0f, 0.0f, 0.0f));
    return false;
  }
  V3f b2 = v3(beta.x / p, beta.y / p, beta.z / p);
  tally_add3(tal, 6, b2, beta);
  beta = b2;
  return true;
}

// Trace photon `pid` from the point light.  The sink receives, in event order,
//   sink.surface(pos, power, dir)  at every diffuse hit (incoming power),
//   sink.volume(pos, power, dir)   at every real medium collision (scattered power),
// with power relative to the photon's initial power and dir the travel direction.
template <class Sink>
PM_HD void trace_photon(const Scene& sc, const TraceParams& tp, uint32_t pid, Sink& sink, Tally9& tal) {
  U4 r0 = rng(tp.seed, pid, 0u, 0u, TAG_EMIT);
  V3f d = sample_sphere(u01(r0.x), u01(r0.y));
  V3f o = v3(sc.light[0], sc.light[1], sc.light[2]);
  V3f beta = v3(1.0f, 1.0f, 1.0f);
  bool in_glass = false;
  const bool media = tp.media_on && sc.med.type != MED_NONE;
  const float alpha = media ? sc.med.sigma_s / (sc.med.sigma_s + sc.med.sigma_a) : 1.0f;
  for (int ev = 0; ev < tp.max_bounces; ++ev) {
    Hit<float> h = intersect(sc, o, d);
    float t_surf = h.kind != HIT_NONE ? h.t : t_exit_front(o, d);
    if (media && !in_glass) {
      float tc;
      if (sample_collision(sc.med, tp.seed, tp.max_delta_iters, pid, ev, o, d, t_surf, &tc)) {
        V3f x = o + d * tc;
        V3f b2 = beta * alpha;
        sink.volume(x, b2, d);
        tally_add3(tal, 0, beta, b2);
        beta = b2;
        U4 rs = rng(tp.seed, pid, (uint32_t)ev, 0u, TAG_SCATTER);
        if (!roulette(u01(rs.z), tp.rr_survive, beta, tal)) return;
        d = normalize(sample_hg(d, sc.med.g, u01(rs.x), u01(rs.y)));
        o = x;
        continue;
      }
    }
    if (h.kind == HIT_NONE) {
      tally_add3(tal, 3, beta, v3(0.0f, 0.0f, 0.0f));
      return;
    }
    V3f x = o + d * h.t;
    V3f n = hit_normal(sc, h, x);
    int mat = hit_material(sc, h.kind, h.id);
    U4 rs = rng(tp.seed, pid, (uint32_t)ev, 0u, TAG_SCATTER);
    bool front = dot(n, d) < 0.0f;
    V3f nf = front ? n : -n;
    if (mat == MAT_DIFFUSE) {
      sink.surface(x, beta, d);
      V3f b2 = mul(beta, hit_albedo<float>(sc, h.kind, h.id));
      tally_add3(tal, 0, beta, b2);
      beta = b2;
      if (!roulette(u01(rs.z), tp.rr_survive, beta, tal)) return;
      d = normalize(sample_cosine(nf, u01(rs.x), u01(rs.y)));
      o = x + nf * tp.eps;
    } else if (mat == MAT_MIRROR) {
      V3f b2 = mul(beta, hit_albedo<float>(sc, h.kind, h.id));
      tally_add3(tal, 0, beta, b2);
      beta = b2;
      d = normalize(reflect(d, nf));
      o = x + nf * tp.eps;
    } else {
      float eta = front ? 1.0f / sc.spheres[h.id].ior : sc.spheres[h.id].ior;
      float ci = -dot(d, nf), ct;
      float R = fresnel(ci, eta, &ct);
      if (u01(rs.w) < R) {
        d = normalize(reflect(d, nf));
        o = x + nf * tp.eps;
      } else {
        d = normalize(refract(d, nf, eta, ci, ct));
        o = x - nf * tp.eps;
        in_glass = front;
      }
    }
  }
  tally_add3(tal, 0, beta, v3(0.0f, 0.0f, 0.0f));  // bounce cap reached
}

// ---------------------------------------------------------------- eye path
struct RenderParams {
  uint32_t seed;
  int width, height, spp;
  float march_step;
  int max_depth;
  float eps;
  float eye[3];
  float focal;
  int media_on;
  int fp64_transmittance;
  int max_steps;
};

PM_HD double exp_(double x) { return ::exp(x); }

// Camera ray for pixel (px, py), sample s: pinhole at `eye` looking +z with
// focal length `focal` (the reference's camera at :933-934 is focal length 1
// at the origin looking +z; here it is normalised and jittered).
template <class T>
PM_HD void camera_ray(const RenderParams& rp, uint32_t px, uint32_t py, uint32_t s, V3<T>* o, V3<T>* d, T* xi) {
  uint32_t pid = py * (uint32_t)rp.width + px;
  U4 rc = rng(rp.seed, pid, s, 0u, TAG_CAMERA);
  T jx = (T)u01(rc.x), jy = (T)u01(rc.y);
  *xi = (T)u01(rc.z);
  T aspect = T(rp.width) / T(rp.height);
  T sx = ((T(px) + jx) / T(rp.width) - T(0.5)) * aspect;
  T sy = -((T(py) + jy) / T(rp.height) - T(0.5));
  *d = normalize(v3<T>(sx, sy, T(rp.focal)));
  *o = v3<T>(T(rp.eye[0]), T(rp.eye[1]), T(rp.eye[2]));
}

// A gather type may answer volume queries one march step late (to merge the
// exact second pass of step j with the first pass of step j + 1): it then
// declares `static constexpr bool kDeferredVolume = true` and provides
//   bool volume_defer(x, wo, w, c) -- queue this point; c = w * value of an
//                                     earlier point when one is ready
//   bool volume_flush(wo, c)       -- the last queued point's w * value (wo:
//                                     the segment's view direction, which is
//                                     also the queued point's)
// The march adds the weighted values in step order either way, so the sum
// is the same as with the immediate g.volume.
template <class G, class = void>
struct DeferredVolume {
  static constexpr bool value = false;
};
template <class G>
struct DeferredVolume<G, decltype((void)G::kDeferredVolume, void())> {
  static constexpr bool value = G::kDeferredVolume;
};

// A gather type that only counts its queries (their number depends on the
// path's geometry, never on its weights) declares
// `static constexpr bool kCountOnly = true`: the march then skips the
// transmittance, and passes zero weights.
template <class G, class = void>
struct CountOnly {
  static constexpr bool value = false;
};
template <class G>
struct CountOnly<G, decltype((void)G::kCountOnly, void())> {
  static constexpr bool value = G::kCountOnly;
};

// Ray-march of the in-scattered radiance over [t0, t1) along o + t d:
// samples at t_j = t0 + (j + xi) * step, each weighted by the transmittance
// from t0 to t_j and the step length.  Homogeneous: T_j = exp(-sigma_t *
// (t_j - t0)).  Density field: the optical depth to t_j over the intervals
// [t0, t_0], [t_0, t_1], ..., [t_{j-1}, t_j], the first by sigma_t(x_0) times
// its length and the others by the trapezoid rule:
// tau_j = sigma_t(x_0) (t_0 - t0) + sum_{0<i<=j} (sigma_t(x_{i-1}) +
// sigma_t(x_i)) / 2 (t_i - t_{i-1}); the segment's own transmittance adds
// sigma_t(x_last) (t1 - t_last) (with no sample, sigma_t at the segment's
// midpoint times its length).  At constant density both forms give
// exp(-sigma_t (t - t0)) at every sample and at t1.  tau is accumulated in TT
// precision.
template <class T, class TT, class G>
PM_HD void march_segment(const Scene& sc, const RenderParams& rp, V3<T> o, V3<T> d, T t0, T t1, T xi,
                         V3<T>& thr, NSum3<T>& L, G& g) {
  const Medium& m = sc.med;
  const T ds = T(rp.march_step);
  const TT st = TT(m.sigma_s) + TT(m.sigma_a);
  TT tau = TT(0), s_last = TT(0);
  T t_last = t0;
  bool sampled = false;
  const V3<T> wo = -d;
  for (int j = 0; j < rp.max_steps; ++j) {
    T t = t0 + (T(j) + xi) * ds;
    if (!(t < t1)) break;
    V3<T> x = o + d * t;
    if constexpr (CountOnly<G>::value) {
      (void)g.volume(x, wo, v3<T>(T(0), T(0), T(0)));
      continue;
    }
    T Tj;
    if (m.type == MED_HOMOGENEOUS) {
      Tj = (T)exp_(-st * (TT)(t - t0));
    } else {
      const TT s = (TT)sigma_t_at(m, x);
      tau = tau + (sampled ? (s_last + s) * TT(0.5) : s) * (TT)(t - t_last);
      Tj = (T)exp_(-tau);
      t_last = t;
      s_last = s;
      sampled = true;
    }
    V3<T> w = thr * (Tj * ds);
    if constexpr (DeferredVolume<G>::value) {
      V3<T> c;
      if (g.volume_defer(x, wo, w, c)) nsum3_add(L, c);
    } else {
      V3<T> S = g.volume(x, wo, w);
      nsum3_add(L, mul(w, S));
    }
  }
  if constexpr (DeferredVolume<G>::value) {
    V3<T> c;
    if (g.volume_flush(wo, c)) nsum3_add(L, c);
  }
  if constexpr (CountOnly<G>::value) return;
  T Tseg;
  if (m.type == MED_HOMOGENEOUS) {
    Tseg = (T)exp_(-st * (TT)(t1 - t0));
  } else {
    if (!sampled) s_last = (TT)sigma_t_at(m, o + d * ((t0 + t1) * T(0.5)));
    Tseg = (T)exp_(-(tau + s_last * (TT)(t1 - t_last)));
  }
  thr = thr * Tseg;
}

// Segments of at least this many march steps in a density field take the
// FP64 transmittance when it is enabled (fp64_transmittance): there the
// optical depth sums the most terms.  Shorter segments, and every segment of
// a homogeneous medium (a closed form, no sum), stay in the frame's precision.
constexpr int kLongSegmentSteps = 32;

// One eye sample.  The gather object provides
//   g.surface(x, n, albedo, w, wo) -> outgoing radiance of a diffuse point toward wo (n faces the viewer)
//   g.volume(x, wo, w)         -> in-scattered source term sigma_s * int p L_i at x toward wo
// where w is the path weight the returned value is multiplied by (throughput,
// and for volume points transmittance times the step length); the sample's
// radiance is the sum of w * value over its queries.
template <class T, class TT, class G>
PM_HD V3<T> render_sample_t(const Scene& sc, const RenderParams& rp, uint32_t px, uint32_t py, uint32_t s, G& g) {
  V3<T> o, d;
  T xi;
  camera_ray(rp, px, py, s, &o, &d, &xi);
  V3<T> thr = v3<T>(T(1), T(1), T(1));
  NSum3<T> L;
  nsum3_init(L);
  bool in_glass = false;
  const bool media = rp.media_on && sc.med.type != MED_NONE;
  const uint32_t pixel = py * (uint32_t)rp.width + px;
  for (int depth = 0; depth < rp.max_depth; ++depth) {
    Hit<T> h = intersect(sc, o, d);
    if (media && !in_glass) {
      T t0 = T(0);
      if (o.z < T(0)) t0 = d.z > T(0) ? -o.z / d.z : T(3.0e38);
      T t1 = h.kind != HIT_NONE ? h.t : (o.z < T(0) ? T(0) : t_exit_front(o, d));
      if (t1 > t0) {
        if constexpr (sizeof(TT) > sizeof(T)) {
          if (sc.med.type == MED_NOISE && t1 - t0 >= T(kLongSegmentSteps) * T(rp.march_step))
            march_segment<T, TT>(sc, rp, o, d, t0, t1, xi, thr, L, g);
          else
            march_segment<T, T>(sc, rp, o, d, t0, t1, xi, thr, L, g);
        } else {
          march_segment<T, TT>(sc, rp, o, d, t0, t1, xi, thr, L, g);
        }
      }
    }
    if (h.kind == HIT_NONE) break;
    V3<T> x = o + d * h.t;
    V3<T> n = hit_normal(sc, h, x);
    bool front = dot(n, d) < T(0);
    V3<T> nf = front ? n : -n;
    int mat = hit_material(sc, h.kind, h.id);
    if (mat == MAT_DIFFUSE) {
      V3<T> Ls = g.surface(x, nf, hit_albedo<T>(sc, h.kind, h.id), thr, -d);
      nsum3_add(L, mul(thr, Ls));
      break;
    } else if (mat == MAT_MIRROR) {
      thr = mul(thr, hit_albedo<T>(sc, h.kind, h.id));
      d = normalize(reflect(d, nf));
      o = x + nf * T(rp.eps);
    } else {
      U4 re = rng(rp.seed, pixel, s, (uint32_t)depth, TAG_EYE);
      T eta = front ? T(1) / T(sc.spheres[h.id].ior) : T(sc.spheres[h.id].ior);
      T ci = -dot(d, nf), ct;
      T R = fresnel(ci, eta, &ct);
      if (T(u01(re.x)) < R) {
        d = normalize(reflect(d, nf));
        o = x + nf * T(rp.eps);
      } else {
        d = normalize(refract(d, nf, eta, ci, ct));
        o = x - nf * T(rp.eps);
        in_glass = front;
      }
    }
  }
  return nsum3_get(L);
}

template <class T, class G>
PM_HD V3<T> render_sample(const Scene& sc, const RenderParams& rp, uint32_t px, uint32_t py, uint32_t s, G& g) {
  if (rp.fp64_transmittance) return render_sample_t<T, double>(sc, rp, px, py, s, g);
  return render_sample_t<T, T>(sc, rp, px, py, s, g);
}

// ---------------------------------------------------------------- estimators
// Density-estimate normalisations (Jensen 2001): surface 1/(pi r^2) with a
// Lambertian BRDF albedo/pi; volume 1/(4/3 pi r^3) with the phase function.
// The cone filter (weight 1 - d/(k r), k = 1.1) is normalised by 1 - 2/(3k).
// Constants are written once, in double, and rounded to T by both sides.
template <class T> PM_HD T surface_norm(T r2, bool cone) {
  T nrm = T(0.318309886183790671537767526745) / (T(3.14159265358979323846264338328) * r2);
  if (cone) nrm = nrm / T(0.393939393939393939393939393939);
  return nrm;
}
template <class T> PM_HD T volume_norm(T r2) {
  return T(1) / (T(4.18879020478639098461685784437) * (r2 * sqrt_(r2)));
}
template <class T> PM_HD T cone_weight(T d2, T r) { return T(1) - sqrt_(d2) / (T(1.1) * r); }

// Display mapping of the FP32 radiance, the reference's frame-write rule
// (photonMappingKernel.cu:1451-1453): c * 255.0 in double, capped at 255 and
// truncated toward zero.  The FP32 radiance is first scaled by the exposure.
// The reference has no lower bound (a negative value converted to unsigned
// char is undefined behaviour); here values below 0, and NaN, give 0.
PM_HD unsigned char tonemap_u8(float v, float exposure) {
  const float c = v * exposure;
  double s = (double)c * 255.0;
  s = s > 255.0 ? 255.0 : s;
  if (!(s > 0.0)) s = 0.0;
  return (unsigned char)(int)s;
}

}  // namespace pm
