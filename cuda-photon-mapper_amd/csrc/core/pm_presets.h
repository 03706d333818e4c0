// pm_presets.h -- host-side configuration: the workload presets and the
// conversion of the C-ABI pm_config into the per-thread Scene/params structs.
//
// Reference counterparts: the compile-time tunables (photonMappingKernel.cu:7-46)
// and the scene globals (:59-98); sphere animation positionObjects (:1380-1404).
#pragma once

#include <string.h>

#include <vector>

#include "../../../include/photon_mapper.h"
#include "pm_transport.h"

namespace pm {

inline void preset_room(pm_config* c) {
  memset(c, 0, sizeof(*c));
  c->seed = 1;
  const float white[3] = {0.7f, 0.7f, 0.7f};
  const float green[3] = {0.07f, 0.7f, 0.07f};
  const float red[3] = {0.7f, 0.07f, 0.07f};
  for (int f = 0; f < 5; ++f)
    for (int k = 0; k < 3; ++k) c->wall_albedo[f][k] = f == 0 ? green[k] : (f == 2 ? red[k] : white[k]);
  c->n_spheres = 2;
  // Lambertian sphere r=0.2 and perfect mirror r=0.25, both resting on the floor.
  pm_sphere* s0 = &c->spheres[0];
  s0->center[0] = 0.28f; s0->center[1] = 0.2f; s0->center[2] = 0.62f; s0->radius = 0.2f;
  s0->albedo[0] = s0->albedo[1] = s0->albedo[2] = 0.7f; s0->material = PM_MAT_DIFFUSE; s0->ior = 1.0f;
  pm_sphere* s1 = &c->spheres[1];
  s1->center[0] = 0.72f; s1->center[1] = 0.25f; s1->center[2] = 0.38f; s1->radius = 0.25f;
  s1->albedo[0] = s1->albedo[1] = s1->albedo[2] = 1.0f; s1->material = PM_MAT_MIRROR; s1->ior = 1.0f;
  c->light_pos[0] = 0.5f; c->light_pos[1] = 0.98f; c->light_pos[2] = 0.5f;
  c->light_power[0] = c->light_power[1] = c->light_power[2] = 100.0f;
  c->medium = PM_MEDIUM_HOMOGENEOUS;
  c->sigma_s = 0.1f; c->sigma_a = 0.01f; c->hg_g = 0.0f;
  c->noise_res = 64; c->noise_freq = 8;
  c->n_emit = 100000; c->cap_surface = 0; c->cap_volume = 0;
  c->max_bounces = 8; c->rr_survive = 0.8f; c->leaf_size = 32;
  c->storage = PM_STORE_FP32;
  c->k_surface = 32; c->k_volume = 32;
  c->rcap_surface = 0.05f; c->rcap_volume = 0.05f;
  c->estimator = PM_EST_BOX;
  c->width = 256; c->height = 256; c->spp = 1;
  c->march_step = 0.01f; c->max_depth = 8;
  c->eye[0] = 0.5f; c->eye[1] = 0.5f; c->eye[2] = -0.8f;
  c->focal = 1.5f; c->exposure = 0.1f;
  c->fp64_transmittance = 0;
  c->media_on = 1;
  c->anim_time = 0.0f;
}

// Returns 0 on success, -1 for an unknown preset.
inline int preset_config(int preset, pm_config* c) {
  preset_room(c);
  switch (preset) {
    case 1:
      return 0;
    case 2:
      c->n_emit = 0;
      c->cap_surface = 1000000u;
      c->cap_volume = 2000000u;
      c->medium = PM_MEDIUM_NOISE;
      c->sigma_s = 0.3f; c->sigma_a = 0.03f; c->hg_g = 0.6f;
      c->k_surface = 64; c->k_volume = 64;
      c->rcap_surface = 0.03f; c->rcap_volume = 0.03f;
      c->leaf_size = 52;  // fastest for the 1080p frame's two-pass gather (48: 1.5 % slower, 32: 4 %, 64: 6 %)
      c->width = 1920; c->height = 1080; c->spp = 4;
      return 0;
    case 3:
      c->storage = PM_STORE_BF16;
      c->fp64_transmittance = 1;
      c->width = 480; c->height = 270;
      return 0;
    case 5:
      // learned-cache scene: preset 1 with the k = 64 gather as the cache's
      // training target and fallback
      c->k_surface = 64;
      c->k_volume = 64;
      return 0;
    default:
      return -1;
  }
}

// Scene `index` of the batched workload: the unit room of preset 1 (same
// walls, light, camera, k = 32, r <= 0.05, 1 spp) with, drawn from Philox
// (key = seed, counter = (index, object, draw, TAG_SCENE)):
//   2 .. 256 spheres, radius U[0.02, 0.2], material 60% Lambertian (albedo
//   U[0.2, 0.8] per channel) / 25% mirror / 15% dielectric (IOR 1.5), centres
//   uniform inside the room, placed without overlapping each other or the
//   light (up to 32 tries, then the sphere is dropped);
//   photons emitted log-uniform in [10^4, 2 10^6];
//   image side from {128, 256, 512, 1024};
//   homogeneous medium, sigma_s U[0, 0.3], sigma_a = sigma_s / 10.
// Returns 0, or -1 for a negative index.
inline int batch_scene_config(uint32_t seed, int index, pm_config* c) {
  if (index < 0) return -1;
  preset_room(c);
  const uint32_t id = (uint32_t)index;
  auto u = [&](uint32_t obj, uint32_t draw) { return u01(rng(seed, id, obj, draw, TAG_SCENE).x); };
  c->seed = seed * 0x9E3779B9u + id + 1u;  // independent photon and camera streams per scene
  const int nsph = 2 + (int)(u(0, 0) * 255.0f);
  c->n_emit = (uint32_t)(1e4f * exp_(u(0, 1) * 5.298317366548036f));  // ln(200)
  const int sides[4] = {128, 256, 512, 1024};
  const int side = sides[(int)(u(0, 2) * 4.0f)];
  c->width = c->height = side;
  c->sigma_s = 0.3f * u(0, 3);
  c->sigma_a = 0.1f * c->sigma_s;
  if (!(c->sigma_s > 0.0f)) c->medium = PM_MEDIUM_NONE;
  c->n_spheres = 0;
  for (int j = 0; j < nsph && c->n_spheres < PM_MAX_SPHERES; ++j) {
    const uint32_t o = 1u + (uint32_t)j;
    pm_sphere sp;
    memset(&sp, 0, sizeof(sp));
    sp.radius = 0.02f + 0.18f * u(o, 0);
    const float m = u(o, 1);
    if (m < 0.6f) {
      sp.material = PM_MAT_DIFFUSE;
      for (int k = 0; k < 3; ++k) sp.albedo[k] = 0.2f + 0.6f * u(o, 2u + (uint32_t)k);
      sp.ior = 1.0f;
    } else {
      sp.material = m < 0.85f ? PM_MAT_MIRROR : PM_MAT_DIELECTRIC;
      sp.albedo[0] = sp.albedo[1] = sp.albedo[2] = 1.0f;
      sp.ior = sp.material == PM_MAT_DIELECTRIC ? 1.5f : 1.0f;
    }
    const float r = sp.radius;
    bool placed = false;
    for (uint32_t a = 0; a < 32 && !placed; ++a) {
      float ctr[3];
      for (int k = 0; k < 3; ++k) ctr[k] = r + (1.0f - 2.0f * r) * u(o, 8u + 3u * a + (uint32_t)k);
      float dx = ctr[0] - c->light_pos[0], dy = ctr[1] - c->light_pos[1], dz = ctr[2] - c->light_pos[2];
      bool ok = (dx * dx + dy * dy) + dz * dz > (r + 0.03f) * (r + 0.03f);
      for (int i = 0; ok && i < c->n_spheres; ++i) {
        const pm_sphere& q = c->spheres[i];
        float ex = ctr[0] - q.center[0], ey = ctr[1] - q.center[1], ez = ctr[2] - q.center[2];
        float rr = r + q.radius + 1e-3f;
        ok = (ex * ex + ey * ey) + ez * ez > rr * rr;
      }
      if (ok) {
        for (int k = 0; k < 3; ++k) sp.center[k] = ctr[k];
        placed = true;
      }
    }
    if (placed) c->spheres[c->n_spheres++] = sp;
  }
  return 0;
}

// Sphere array with the legacy animation applied: each sphere's centre moves
// on a small horizontal circle of radius 0.03 with phase anim_time + i*pi
// (the reference moves spheres 0 and 1 along sin/cos paths, :1380-1392).
inline std::vector<Sphere> build_spheres(const pm_config& c) {
  std::vector<Sphere> v((size_t)c.n_spheres);
  for (int i = 0; i < c.n_spheres; ++i) {
    const pm_sphere& s = c.spheres[i];
    Sphere& o = v[(size_t)i];
    memset(&o, 0, sizeof(o));
    o.cx = s.center[0]; o.cy = s.center[1]; o.cz = s.center[2]; o.r = s.radius;
    o.ar = s.albedo[0]; o.ag = s.albedo[1]; o.ab = s.albedo[2];
    o.mat = s.material; o.ior = s.ior;
    if (c.anim_time != 0.0f) {
      float turns = c.anim_time * 0.159154943091895336f + 0.5f * (float)i;
      turns = turns - floor_(turns);
      float sn, cs;
      sincos_turns(turns, &sn, &cs);
      o.cx = o.cx + 0.03f * sn;
      o.cz = o.cz + 0.03f * cs;
    }
  }
  return v;
}

inline std::vector<float> build_density(const pm_config& c) {
  std::vector<float> d;
  if (c.medium != PM_MEDIUM_NOISE) return d;
  const int R = c.noise_res;
  d.resize((size_t)R * R * R);
  for (int k = 0; k < R; ++k)
    for (int j = 0; j < R; ++j)
      for (int i = 0; i < R; ++i)
        d[((size_t)k * R + j) * R + i] = noise_vertex(c.seed, R, c.noise_freq, i, j, k);
  return d;
}

inline Scene make_scene(const pm_config& c, const Sphere* spheres, const float* density) {
  Scene s;
  memset(&s, 0, sizeof(s));
  for (int f = 0; f < 5; ++f)
    for (int k = 0; k < 3; ++k) s.wall_albedo[f][k] = c.wall_albedo[f][k];
  s.nspheres = c.n_spheres;
  s.spheres = spheres;
  for (int k = 0; k < 3; ++k) {
    s.light[k] = c.light_pos[k];
    s.light_power[k] = c.light_power[k];
  }
  s.med.type = c.medium;
  s.med.sigma_s = c.sigma_s;
  s.med.sigma_a = c.sigma_a;
  s.med.g = c.hg_g;
  s.med.res = c.noise_res;
  s.med.majorant = c.sigma_s + c.sigma_a;  // density <= 1
  s.med.density = density;
  return s;
}

inline TraceParams trace_params(const pm_config& c) {
  TraceParams t;
  t.seed = c.seed;
  t.max_bounces = c.max_bounces;
  t.rr_survive = c.rr_survive;
  t.eps = 1e-4f;
  t.max_delta_iters = 256;
  t.media_on = c.media_on && c.medium != PM_MEDIUM_NONE;
  return t;
}

inline RenderParams render_params(const pm_config& c) {
  RenderParams r;
  r.seed = c.seed;
  r.width = c.width;
  r.height = c.height;
  r.spp = c.spp;
  r.march_step = c.march_step;
  r.max_depth = c.max_depth;
  r.eps = 1e-4f;
  for (int k = 0; k < 3; ++k) r.eye[k] = c.eye[k];
  r.focal = c.focal;
  r.media_on = c.media_on && c.medium != PM_MEDIUM_NONE;
  r.fp64_transmittance = c.fp64_transmittance;
  r.max_steps = 1 << 20;
  return r;
}

inline bool finite_f(float v) { return __builtin_isfinite(v); }
inline bool unit_f(float v) { return v >= 0.0f && v <= 1.0f; }  // false for NaN

// Validates a configuration; returns 0 or PM_ERR_INVALID.  Every float that
// reaches the kernels must be finite (NaN or Inf in a position, power or
// coefficient would propagate silently into the photon maps and the frame).
inline int validate_config(const pm_config& c) {
  if (c.n_spheres < 0 || c.n_spheres > PM_MAX_SPHERES) return PM_ERR_INVALID;
  for (int f = 0; f < 5; ++f)
    for (int k = 0; k < 3; ++k)
      if (!unit_f(c.wall_albedo[f][k])) return PM_ERR_INVALID;
  for (int k = 0; k < 3; ++k) {
    if (!finite_f(c.light_pos[k]) || !finite_f(c.eye[k])) return PM_ERR_INVALID;
    if (!finite_f(c.light_power[k]) || c.light_power[k] < 0.0f) return PM_ERR_INVALID;
  }
  if (!finite_f(c.sigma_s) || !finite_f(c.sigma_a) || c.sigma_s < 0.0f || c.sigma_a < 0.0f) return PM_ERR_INVALID;
  if (!finite_f(c.hg_g) || !finite_f(c.anim_time)) return PM_ERR_INVALID;
  if (!finite_f(c.focal) || !(c.focal > 0.0f) || !finite_f(c.exposure) || c.exposure < 0.0f) return PM_ERR_INVALID;
  if (!finite_f(c.march_step) || !finite_f(c.rcap_surface) || !finite_f(c.rcap_volume)) return PM_ERR_INVALID;
  if (c.max_bounces < 1 || c.max_bounces > 64) return PM_ERR_INVALID;
  if (!(c.rr_survive > 0.0f) || c.rr_survive > 1.0f) return PM_ERR_INVALID;
  if (c.leaf_size < 1 || c.leaf_size > 64) return PM_ERR_INVALID;
  if (c.k_surface < 1 || c.k_surface > 64 || c.k_volume < 1 || c.k_volume > 64) return PM_ERR_INVALID;
  if (!(c.rcap_surface > 0.0f) || !(c.rcap_volume > 0.0f)) return PM_ERR_INVALID;
  if (c.width < 1 || c.height < 1 || c.spp < 1 || c.width > 16384 || c.height > 16384 || c.spp > 1024)
    return PM_ERR_INVALID;
  if (!(c.march_step > 0.0f) || c.max_depth < 1 || c.max_depth > 64) return PM_ERR_INVALID;
  if (c.medium < PM_MEDIUM_NONE || c.medium > PM_MEDIUM_NOISE) return PM_ERR_INVALID;
  if (c.medium == PM_MEDIUM_NOISE && (c.noise_res < 2 || c.noise_res > 512 || c.noise_freq < 1)) return PM_ERR_INVALID;
  if (c.medium != PM_MEDIUM_NONE && !(c.sigma_s + c.sigma_a > 0.0f)) return PM_ERR_INVALID;
  if (c.hg_g <= -1.0f || c.hg_g >= 1.0f) return PM_ERR_INVALID;
  if (c.n_emit == 0 && (c.cap_surface == 0 || c.cap_volume == 0) && c.media_on && c.medium != PM_MEDIUM_NONE)
    return PM_ERR_INVALID;
  if (c.n_emit == 0 && c.cap_surface == 0) return PM_ERR_INVALID;
  for (int i = 0; i < c.n_spheres; ++i) {
    const pm_sphere& s = c.spheres[i];
    if (!finite_f(s.radius) || !(s.radius > 0.0f) || s.material < 0 || s.material > 2) return PM_ERR_INVALID;
    for (int k = 0; k < 3; ++k)
      if (!finite_f(s.center[k]) || !unit_f(s.albedo[k])) return PM_ERR_INVALID;
    if (!finite_f(s.ior)) return PM_ERR_INVALID;
    if (s.material == PM_MAT_DIELECTRIC && !(s.ior > 1.0f)) return PM_ERR_INVALID;
  }
  if (c.storage != PM_STORE_FP32 && c.storage != PM_STORE_BF16) return PM_ERR_INVALID;
  return 0;
}

}  // namespace pm
