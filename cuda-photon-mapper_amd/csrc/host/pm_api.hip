// pm_api.hip -- the C ABI of libpm.so (include/photon_mapper.h): the library
// context, the pm_* entry points and the reference's three launchers.
//
// Reference boundary replaced: photonMappingKernel.cu:1523-1582
// (launch_emit_photons_kernel, launch_photon_mapping_kernel,
// launch_init_random_numbers_kernel) with their error check at :49-55, as
// called from simplePBO.cpp:126,141,177.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "pm_context.h"

using namespace pm;

namespace pm {
void cache_release(CacheState& k);
char g_last_err[256] = "";
void set_last_hip(hipError_t e, const char* what) {
  snprintf(g_last_err, sizeof(g_last_err), "%s: %s", what, hipGetErrorString(e));
}
}  // namespace pm

namespace {

int setup_scene(pm_context* c) {
  c->spheres_h = build_spheres(c->cfg);
  HIPCHK(c->spheres_d.ensure(sizeof(Sphere) * std::max<size_t>(1, c->spheres_h.size())));
  if (!c->spheres_h.empty())
    HIPCHK(hipMemcpy(c->spheres_d.p, c->spheres_h.data(), sizeof(Sphere) * c->spheres_h.size(), hipMemcpyHostToDevice));
  std::vector<float> dens = build_density(c->cfg);
  if (!dens.empty()) {
    HIPCHK(c->density_d.ensure(sizeof(float) * dens.size()));
    HIPCHK(hipMemcpy(c->density_d.p, dens.data(), sizeof(float) * dens.size(), hipMemcpyHostToDevice));
  }
  c->scene = make_scene(c->cfg, c->spheres_d.as<Sphere>(), dens.empty() ? nullptr : c->density_d.as<float>());
  c->tp = trace_params(c->cfg);
  c->rp = render_params(c->cfg);
  return PM_OK;
}

// Grow a u32 array to `words`, keeping the first `keep` words.
int grow_keep(DevBuf& b, size_t words, size_t keep, hipStream_t st) {
  if (b.p && b.bytes >= words * 4) return PM_OK;
  DevBuf nb;
  HIPCHK(nb.ensure(words * 4));
  if (keep && b.p) HIPCHK(hipMemcpyAsync(nb.p, b.p, keep * 4, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  b.release();
  b = nb;
  return PM_OK;
}

int alloc_map(MapBufs& mb, uint32_t n, int bf16, int L) {
  HIPCHK(mb.x.ensure(4ull * std::max<uint32_t>(n, 1)));
  HIPCHK(mb.y.ensure(4ull * std::max<uint32_t>(n, 1)));
  HIPCHK(mb.z.ensure(4ull * std::max<uint32_t>(n, 1)));
  HIPCHK(mb.payload.ensure((bf16 ? 16ull : 32ull) * std::max<uint32_t>(n, 1)));
  HIPCHK(mb.tpayload.ensure((bf16 ? 16ull : 32ull) * std::max<uint32_t>(n, 1)));
  HIPCHK(mb.tpts.ensure(16ull * ((uint64_t)n + kTptsPad)));
  uint32_t nl = kd_num_leaves(n, (uint32_t)L);
  HIPCHK(mb.nodes.ensure(sizeof(KdNode) * std::max<uint32_t>(nl, 1)));
  HIPCHK(mb.nrange.ensure(sizeof(uint2) * std::max<uint32_t>(nl, 1)));
  HIPCHK(mb.bmin.ensure(sizeof(float4) * std::max<uint32_t>(2 * nl, 1)));
  HIPCHK(mb.bmax.ensure(sizeof(float4) * std::max<uint32_t>(2 * nl, 1)));
  DevMap& m = mb.m;
  memset(&m, 0, sizeof(m));
  m.n = n;
  m.cap = n;
  m.bf16 = bf16;
  m.x = mb.x.as<float>();
  m.y = mb.y.as<float>();
  m.z = mb.z.as<float>();
  m.payload = mb.payload.p;
  m.L = (uint32_t)L;
  m.nl = nl;
  m.tpts = mb.tpts.as<float4>();
  m.nodes = mb.nodes.as<KdNode>();
  m.nrange = mb.nrange.as<uint2>();
  m.bmin = mb.bmin.as<float4>();
  m.bmax = mb.bmax.as<float4>();
  m.tpayload = mb.tpayload.p;
  m.built = false;
  return PM_OK;
}

double elapsed_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return -1.0;
  return (double)ms;
}

int do_emit(pm_context* c) {
  hipStream_t st = c->stream;
  const bool media = c->tp.media_on != 0;
  const bool fixed = c->cfg.n_emit > 0;
  const uint64_t cap_s = c->cfg.cap_surface ? c->cfg.cap_surface : UINT64_MAX;
  const uint64_t cap_v = (c->cfg.cap_volume && media) ? c->cfg.cap_volume : UINT64_MAX;
  HIPCHK(hipEventRecord(c->e0, st));
  HIPCHK(c->small.ensure(64));
  uint32_t* small = c->small.as<uint32_t>();
  // Record counts, offsets and totals are 32-bit.  A photon stores at most
  // max_bounces records, so at most UINT32_MAX / max_bounces photons are
  // traced: no total can wrap, and k_find_cap sees each offset range once.
  const uint64_t max_photons = std::min<uint64_t>(1ull << 31, 0xffffffffull / (uint64_t)c->cfg.max_bounces);
  if (fixed && c->cfg.n_emit > max_photons) return PM_ERR_CAPACITY;
  uint32_t N = fixed ? c->cfg.n_emit
                     : (uint32_t)std::min<uint64_t>(std::min<uint64_t>(std::max<uint64_t>(cap_s, 1u << 16), 1u << 30),
                                                    max_photons);
  // caps mode: start from the photon count the last emission needed (its
  // N_total).  The records are those of the first N_total photons whatever N
  // the loop reaches, so a larger start only traces more photons through the
  // count pass, and the same scene needs one pass instead of a doubling
  // sequence.
  if (!fixed && c->caps_N > N) N = (uint32_t)std::min<uint64_t>(c->caps_N, max_photons);
  uint32_t done = 0;
  uint32_t tot[2] = {0, 0};
  while (true) {
    int r;
    if ((r = grow_keep(c->cnt_s, N, done, st)) != PM_OK) return r;
    if ((r = grow_keep(c->cnt_v, N, done, st)) != PM_OK) return r;
    HIPCHK(c->off_s.ensure(4ull * N));
    HIPCHK(c->off_v.ensure(4ull * N));
    HIPCHK(c->scratch.ensure(4ull * scan_temp_words(N)));
    HIPCHK(launch_trace_count(c->scene, c->tp, done, N - done, c->cnt_s.as<uint32_t>(), c->cnt_v.as<uint32_t>(), st));
    done = N;
    HIPCHK(exclusive_scan(c->cnt_s.as<uint32_t>(), c->off_s.as<uint32_t>(), N, c->scratch.as<uint32_t>(), small, st));
    HIPCHK(exclusive_scan(c->cnt_v.as<uint32_t>(), c->off_v.as<uint32_t>(), N, c->scratch.as<uint32_t>(), small + 1,
                          st));
    HIPCHK(hipMemcpyAsync(tot, small, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (fixed) break;
    if (tot[0] >= cap_s && (!media || tot[1] >= cap_v)) break;
    if (N >= max_photons) return PM_ERR_CAPACITY;
    N = (uint32_t)std::min<uint64_t>((uint64_t)N * 2u, max_photons);
  }
  uint32_t n_s = (uint32_t)std::min<uint64_t>(tot[0], cap_s);
  uint32_t n_v = media ? (uint32_t)std::min<uint64_t>(tot[1], cap_v) : 0u;
  uint32_t N_s = N, N_v = N, N_total = N;
  if (!fixed) {
    uint32_t init[2] = {0xffffffffu, 0xffffffffu};
    HIPCHK(hipMemcpyAsync(small + 2, init, 8, hipMemcpyHostToDevice, st));
    HIPCHK(launch_find_cap(c->off_s.as<uint32_t>(), c->cnt_s.as<uint32_t>(), N, (uint32_t)cap_s, small + 2, st));
    if (media)
      HIPCHK(launch_find_cap(c->off_v.as<uint32_t>(), c->cnt_v.as<uint32_t>(), N, (uint32_t)cap_v, small + 3, st));
    uint32_t fc[2];
    HIPCHK(hipMemcpyAsync(fc, small + 2, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    N_s = fc[0];
    N_v = media ? fc[1] : 0u;
    if (N_s == 0xffffffffu || (media && N_v == 0xffffffffu)) return PM_ERR_STATE;
    N_total = media ? std::max(N_s, N_v) : N_s;
    if (!media) N_v = N_total;
    // the next emission starts its count pass at the photon that met both caps
    c->caps_N = N_total;
  }
  int r;
  const int bf16 = c->cfg.storage == PM_STORE_BF16;
  if ((r = alloc_map(c->ms, n_s, bf16, c->cfg.leaf_size)) != PM_OK) return r;
  if ((r = alloc_map(c->mv, n_v, bf16, c->cfg.leaf_size)) != PM_OK) return r;
  c->ms.m.n_emit = N_s;
  c->mv.m.n_emit = N_v;
  uint32_t nb = trace_write_blocks(N_total);
  HIPCHK(c->tally_part.ensure(sizeof(double) * 9 * std::max<uint32_t>(nb, 1)));
  HIPCHK(c->tally9.ensure(sizeof(double) * 9));
  HIPCHK(launch_trace_write(c->scene, c->tp, N_total, c->off_s.as<uint32_t>(), c->off_v.as<uint32_t>(), c->ms.m,
                            c->mv.m, c->tally_part.as<double>(), st));
  HIPCHK(launch_tally_final(c->tally_part.as<double>(), nb, c->tally9.as<double>(), st));
  HIPCHK(hipMemcpyAsync(c->tally_h, c->tally9.p, sizeof(double) * 9, hipMemcpyDeviceToHost, st));
  HIPCHK(hipEventRecord(c->e1, st));
  HIPCHK(hipStreamSynchronize(st));
  c->n_emit_total = N_total;
  c->stats.ms_emit = elapsed_ms(c->e0, c->e1);
  c->stats.n_surface = n_s;
  c->stats.n_volume = n_v;
  c->stats.n_emit_surface = N_s;
  c->stats.n_emit_volume = N_v;
  c->stats.n_emit_total = N_total;
  for (int k = 0; k < 3; ++k) {
    double w = (double)c->cfg.light_power[k] / (double)N_total;
    c->stats.emitted[k] = (double)c->cfg.light_power[k];
    c->stats.absorbed[k] = c->tally_h[k] * w;
    c->stats.escaped[k] = c->tally_h[3 + k] * w;
    c->stats.rr_gain[k] = c->tally_h[6 + k] * w;
  }
  c->emitted = true;
  c->built = false;
  c->vox_ready = false;
  return PM_OK;
}

}  // namespace

// Host shape tables of the canonical tree (see core/pm_kdtree.h).
void pm::shape_tables(uint32_t n, uint32_t L, std::vector<uint32_t>& lo, std::vector<uint32_t>& hi,
                  std::vector<uint32_t>& lc, std::vector<uint32_t>& lb) {
  uint32_t nl = kd_num_leaves(n, L);
  size_t ni = nl > 0 ? nl - 1 : 0;
  lo.assign(ni, 0); hi.assign(ni, 0); lc.assign(ni, 0); lb.assign(ni, 0);
  if (ni == 0) return;
  lo[0] = 0;
  hi[0] = n;
  for (size_t h = 0; h < ni; ++h) {
    lc[h] = L * kd_leaves_under(2u * (uint32_t)h + 1u, nl);
    size_t l = 2 * h + 1, r = 2 * h + 2;
    if (l < ni) { lo[l] = lo[h]; hi[l] = lo[h] + lc[h]; }
    if (r < ni) { lo[r] = lo[h] + lc[h]; hi[r] = hi[h]; }
  }
  for (uint64_t f = 0, w = 1; f < ni; f = 2 * f + 1, w *= 2) {
    uint32_t run = 0;
    for (uint64_t h = f; h < std::min<uint64_t>(f + w, ni); ++h) {
      lb[h] = run;
      run += lc[h];
    }
  }
}

namespace {

// Builds mb's tree on stream st with temp as its scratch.
int build_map(pm_context* c, MapBufs& mb, DevBuf& temp, hipStream_t st) {
  DevMap& m = mb.m;
  if (!mb.shape.p || mb.shape_n != m.n || mb.shape_L != m.L) {
    // the shape tables depend on (n, L) only: computed and uploaded once per size
    std::vector<uint32_t> lo, hi, lc, lb;
    shape_tables(m.n, m.L, lo, hi, lc, lb);
    const size_t stride = kd_shape_table_words(m.n, m.L) / 4;
    std::vector<uint32_t> t(4 * stride, 0u);
    for (size_t i = 0; i < lo.size(); ++i) {
      t[i] = lo[i];
      t[stride + i] = hi[i];
      t[2 * stride + i] = lc[i];
      t[3 * stride + i] = lb[i];
    }
    HIPCHK(hipStreamSynchronize(st));  // no build still reads the old tables
    HIPCHK(mb.shape.ensure(4ull * t.size()));
    HIPCHK(hipMemcpy(mb.shape.p, t.data(), 4ull * t.size(), hipMemcpyHostToDevice));
    mb.shape_n = m.n;
    mb.shape_L = m.L;
  }
  HIPCHK(temp.ensure(4ull * kdbuild_temp_words(m.n, m.L)));
  HIPCHK(kd_build_impl(m, temp.as<uint32_t>(), mb.shape.as<uint32_t>(), st));
  return PM_OK;
}

GatherMap gather_map(const pm_context* c, const DevMap& m, int k, float rcap);

}  // namespace

GatherMap pm::make_gather_map(const pm_context* c, const DevMap& m, int k, float rcap) {
  return gather_map(c, m, k, rcap);
}

namespace {

GatherMap gather_map(const pm_context* c, const DevMap& m, int k, float rcap) {
  GatherMap g;
  memset(&g, 0, sizeof(g));
  g.tpts = m.tpts;
  g.nodes = m.nodes;
  g.nrange = m.nrange;
  g.bmin = m.bmin;
  g.bmax = m.bmax;
  g.payload = m.payload;
  g.tpayload = m.tpayload;
  g.n = m.n;
  g.nl = m.nl;
  g.bf16 = m.bf16;
  g.k = k;
  g.rcap2 = rcap * rcap;
  for (int j = 0; j < 3; ++j) g.scale[j] = c->cfg.light_power[j] / (float)(m.n_emit ? m.n_emit : 1u);
  return g;
}

pm_context* g_legacy = nullptr;

}  // namespace

int pm::records_to_host(const DevMap& m, float* pos, float* power, float* dir) {
  const uint32_t n = m.n;
  if (n == 0) return PM_OK;
  std::vector<float> x(n), y(n), z(n);
  HIPCHK(hipMemcpy(x.data(), m.x, 4ull * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(y.data(), m.y, 4ull * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(z.data(), m.z, 4ull * n, hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < n; ++i) {
    pos[3 * i] = x[i]; pos[3 * i + 1] = y[i]; pos[3 * i + 2] = z[i];
  }
  if (m.bf16) {
    std::vector<uint16_t> p(8ull * n);
    HIPCHK(hipMemcpy(p.data(), m.payload, 16ull * n, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i) {
      for (int k = 0; k < 3; ++k) {
        power[3 * i + k] = bf16_to_f32(p[8ull * i + k]);
        dir[3 * i + k] = bf16_to_f32(p[8ull * i + 3 + k]);
      }
    }
  } else {
    std::vector<float> p(8ull * n);
    HIPCHK(hipMemcpy(p.data(), m.payload, 32ull * n, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i) {
      for (int k = 0; k < 3; ++k) {
        power[3 * i + k] = p[8ull * i + k];
        dir[3 * i + k] = p[8ull * i + 3 + k];
      }
    }
  }
  return PM_OK;
}

int pm::tree_to_host(const DevMap& m, uint32_t* perm, float* split, uint32_t* axis) {
  const uint32_t n = m.n, ni = m.nl > 0 ? m.nl - 1 : 0;
  if (n == 0) return PM_OK;
  std::vector<float4> t(n);
  HIPCHK(hipMemcpy(t.data(), m.tpts, 16ull * n, hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < n; ++i) perm[i] = f2u(t[i].w);
  if (ni) {
    if (!split || !axis) return PM_ERR_INVALID;
    std::vector<KdNode> nd(ni);
    HIPCHK(hipMemcpy(nd.data(), m.nodes, sizeof(KdNode) * ni, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < ni; ++i) {
      split[i] = nd[i].split;
      axis[i] = (uint32_t)kd_axis(nd[i].meta);
    }
  }
  return PM_OK;
}

namespace {

void legacy_fail(const char* what, int st) {
  fprintf(stderr, "photon mapper error: %s failed: %s (%s).\n", what, pm_status_string(st), g_last_err);
  exit(EXIT_FAILURE);
}

}  // namespace

extern "C" {

const char* pm_status_string(int s) {
  switch (s) {
    case PM_OK: return "ok";
    case PM_ERR_INVALID: return "invalid argument";
    case PM_ERR_HIP: return "HIP runtime error";
    case PM_ERR_STATE: return "call order violated";
    case PM_ERR_CAPACITY: return "capacity exceeded";
    case PM_ERR_NOMEM: return "out of memory";
    default: return "unknown status";
  }
}

int pm_config_preset(int preset, pm_config* out) {
  if (!out) return PM_ERR_INVALID;
  return preset_config(preset, out) == 0 ? PM_OK : PM_ERR_INVALID;
}

int pm_create(const pm_config* cfg, pm_context** out) {
  if (!cfg || !out) return PM_ERR_INVALID;
  if (validate_config(*cfg) != 0) return PM_ERR_INVALID;
  pm_context* c = new pm_context();
  c->cfg = *cfg;
  memset(&c->stats, 0, sizeof(c->stats));
  hipError_t e = hipGetDevice(&c->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreate(&c->e0);
  if (e == hipSuccess) e = hipEventCreate(&c->e1);
  if (e == hipSuccess) e = hipEventCreate(&c->e2);
  if (e != hipSuccess) {
    set_last_hip(e, "pm_create");
    pm_destroy(c);
    return PM_ERR_HIP;
  }
  int r = setup_scene(c);
  if (r != PM_OK) {
    pm_destroy(c);
    return r;
  }
  *out = c;
  return PM_OK;
}

int pm_destroy(pm_context* c) {
  if (!c) return PM_ERR_INVALID;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  cache_release(c->cache);
  c->ms.release();
  c->mv.release();
  for (DevBuf* b : {&c->spheres_d, &c->density_d, &c->cnt_s, &c->cnt_v, &c->off_s, &c->off_v, &c->scratch,
                    &c->scratch2, &c->tally_part, &c->tally9, &c->small, &c->counters, &c->vox_grid_s, &c->vox_grid_v,
                    &c->vox_sat_s, &c->vox_sat_v})
    b->release();
  if (c->e0) (void)hipEventDestroy(c->e0);
  if (c->e1) (void)hipEventDestroy(c->e1);
  if (c->e2) (void)hipEventDestroy(c->e2);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PM_OK;
}

int pm_emit(pm_context* c) {
  if (!c) return PM_ERR_INVALID;
  return do_emit(c);
}

int pm_build(pm_context* c) {
  if (!c) return PM_ERR_INVALID;
  if (!c->emitted) return PM_ERR_STATE;
  HIPCHK(hipEventRecord(c->e0, c->stream));
  // the two maps' builds are independent and each is a chain of small,
  // latency-bound level passes: the volume map's runs on a second stream
  // with its own scratch, and the first stream waits for it
  HIPCHK(hipEventRecord(c->ev_fork, c->stream));
  HIPCHK(hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
  int r;
  if ((r = build_map(c, c->ms, c->scratch, c->stream)) != PM_OK) return r;
  if ((r = build_map(c, c->mv, c->scratch2, c->stream2)) != PM_OK) return r;
  HIPCHK(hipEventRecord(c->ev_join, c->stream2));
  HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  HIPCHK(hipEventRecord(c->e1, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->stats.ms_build = elapsed_ms(c->e0, c->e1);
  c->built = true;
  return PM_OK;
}

int pm_set_render(pm_context* c, int width, int height, int spp, int k_s, int k_v, float rcap_s, float rcap_v) {
  if (!c) return PM_ERR_INVALID;
  pm_config t = c->cfg;
  t.width = width; t.height = height; t.spp = spp;
  t.k_surface = k_s; t.k_volume = k_v;
  t.rcap_surface = rcap_s; t.rcap_volume = rcap_v;
  if (validate_config(t) != 0) return PM_ERR_INVALID;
  c->cfg = t;
  c->rp = render_params(t);
  return PM_OK;
}

}  // extern "C"

namespace {

// The render behind pm_render_crop and pm_render_crop_noalloc.  `counters`
// null: the context's own query counters (allocated on first use, cleared
// with hipMemsetAsync); otherwise two caller-side counters, cleared by a fill
// kernel, and nothing is allocated.
int render_crop_impl(pm_context* c, int x0, int y0, int w, int h, float* rgb_dev, pm_uchar4* out_dev,
                     unsigned long long* counters) {
  if (!c) return PM_ERR_INVALID;
  // the voxel-grid estimate needs the records only, the k-NN gather their trees
  if (c->voxel_grid ? !c->emitted : !c->built) return PM_ERR_STATE;
  if (w <= 0 || h <= 0 || x0 < 0 || y0 < 0 || x0 + w > c->cfg.width || y0 + h > c->cfg.height) return PM_ERR_INVALID;
  // the caller that allocates nothing needs grids built by an earlier render
  if (c->voxel_grid && counters && !c->vox_ready) return PM_ERR_STATE;
  RenderLaunch r;
  memset(&r, 0, sizeof(r));
  r.sc = c->scene;
  r.rp = c->rp;
  if (!c->render_media) r.rp.media_on = 0;
  r.gs = gather_map(c, c->ms.m, c->cfg.k_surface, c->cfg.rcap_surface);
  r.gv = gather_map(c, c->mv.m, c->cfg.k_volume, c->cfg.rcap_volume);
  r.estimator = c->cfg.estimator;
  r.exposure = c->cfg.exposure;
  r.x0 = x0; r.y0 = y0; r.w = w; r.h = h;
  r.rgb = rgb_dev;
  r.u8 = (uchar4*)out_dev;
  r.blocks_per_cu = c->blocks_per_cu;
  r.gather_mode = c->gather_mode;
  r.tile_regions = c->tile_regions;
  r.sample_lanes = c->sample_lanes;
  r.top_levels = c->top_levels;
  if (counters) {
    r.counters = counters;
    HIPCHK(launch_fill_u32(reinterpret_cast<uint32_t*>(counters), 0u, 4, c->stream));
  } else {
    HIPCHK(c->counters.ensure(16));
    r.counters = c->counters.as<unsigned long long>();
    HIPCHK(hipMemsetAsync(c->counters.p, 0, 16, c->stream));
  }
  if (c->voxel_grid) {
    const size_t grid_bytes = 8ull * 3 * kVox * kVox * kVox, sat_bytes = 8ull * 3 * kSat * kSat * kSat;
    if (!c->vox_ready) {
      HIPCHK(c->vox_grid_s.ensure(grid_bytes));
      HIPCHK(c->vox_grid_v.ensure(grid_bytes));
      HIPCHK(c->vox_sat_s.ensure(sat_bytes));
      HIPCHK(c->vox_sat_v.ensure(sat_bytes));
      HIPCHK(launch_voxel_grids(c->ms.m, c->mv.m, c->vox_grid_s.as<unsigned long long>(),
                                c->vox_grid_v.as<unsigned long long>(), c->vox_sat_s.as<unsigned long long>(),
                                c->vox_sat_v.as<unsigned long long>(), c->stream));
      c->vox_ready = true;
    }
    VoxelGrids G;
    G.sat_s = c->vox_sat_s.as<uint64_t>();
    G.sat_v = c->vox_sat_v.as<uint64_t>();
    for (int j = 0; j < 3; ++j) {
      G.scale_s[j] = c->cfg.light_power[j] / (float)(c->ms.m.n_emit ? c->ms.m.n_emit : 1u);
      G.scale_v[j] = c->cfg.light_power[j] / (float)(c->mv.m.n_emit ? c->mv.m.n_emit : 1u);
    }
    G.interp = c->voxel_grid == 2 ? 1 : 0;
    HIPCHK(hipEventRecord(c->e0, c->stream));
    HIPCHK(launch_render_voxel(r, G, c->stream));
  } else {
  HIPCHK(hipEventRecord(c->e0, c->stream));
  if (r.gather_mode == 5) {
    // auto: the two-pass gather (march passes merged, first passes counted)
    // pays off when the frame keeps many more waves in flight than the 8-byte
    // heaps allow and the maps are large; small frames or maps run the
    // one-pass gather
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device);
    const uint64_t tiles = (uint64_t)((w + 7) / 8) * (uint64_t)((h + 7) / 8);
    const uint64_t recs = (uint64_t)c->ms.m.n + (uint64_t)c->mv.m.n;
    r.gather_mode = tiles >= 16ull * (uint64_t)cus && recs >= (1ull << 20) ? 7 : 1;
  }
  switch (r.gather_mode) {
    case 0: HIPCHK(launch_render(r, c->stream)); break;
    case 1: HIPCHK(launch_render_lane(r, c->stream)); break;
    case 4: HIPCHK(launch_render_lane4(r, c->stream)); break;
    case 7: HIPCHK(launch_render_lane7(r, c->stream)); break;
    default: HIPCHK(launch_render_lane2(r, c->stream)); break;
  }
  }
  HIPCHK(hipEventRecord(c->e1, c->stream));
  unsigned long long q[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(q, r.counters, 16, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->stats.ms_render = elapsed_ms(c->e0, c->e1);
  c->stats.ms_gather_kernel = c->stats.ms_render;
  c->stats.surface_queries = q[0];
  c->stats.volume_queries = q[1];
  return PM_OK;
}

// Scratch of pm_render_crop_noalloc: the render's two query counters
// (surface, volume), 16 bytes in a device array that exists from library load
// to process end, one copy per GPU.  Every call on a GPU shares that GPU's
// copy: calls that overlap on one GPU (two host threads, two contexts) race
// on it.  Each call clears it before its render.
constexpr int kScratchCounters = 2;
static __device__ __align__(256) unsigned long long scratch[kScratchCounters];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned long long* g_scratch_dev[kMaxScratchDevices] = {};

// The device address of `scratch` on `device`, looked up once per device with
// that device current (each GPU holds its own copy at its own address).
int scratch_on(int device, unsigned long long** out) {
  if (device < 0 || device >= kMaxScratchDevices) return PM_ERR_INVALID;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    int cur = -1;
    HIPCHK(hipGetDevice(&cur));
    if (cur != device) HIPCHK(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) {
      set_last_hip(e, "hipGetSymbolAddress(scratch)");
      return PM_ERR_HIP;
    }
    g_scratch_dev[device] = static_cast<unsigned long long*>(p);
  }
  *out = g_scratch_dev[device];
  return PM_OK;
}

}  // namespace

extern "C" {

int pm_render_crop(pm_context* c, int x0, int y0, int w, int h, float* rgb_dev, pm_uchar4* out_dev) {
  return render_crop_impl(c, x0, y0, w, h, rgb_dev, out_dev, nullptr);
}

int pm_render_crop_noalloc(pm_context* c, int x0, int y0, int w, int h, float* rgb_dev, pm_uchar4* out_dev) {
  if (!c) return PM_ERR_INVALID;
  unsigned long long* counters = nullptr;
  int r = scratch_on(c->device, &counters);
  if (r != PM_OK) return r;
  return render_crop_impl(c, x0, y0, w, h, rgb_dev, out_dev, counters);
}

int pm_render(pm_context* c, float* rgb_dev, pm_uchar4* out_dev) {
  if (!c) return PM_ERR_INVALID;
  return pm_render_crop(c, 0, 0, c->cfg.width, c->cfg.height, rgb_dev, out_dev);
}

int pm_set_option(pm_context* c, int option, int value) {
  if (!c) return PM_ERR_INVALID;
  switch (option) {
    case PM_OPT_GATHER_MODE:
      // 3 and 6 were the 8-point-batch and uncounted forms of modes 4 and 7
      if (value < 0 || value > 7 || value == 3 || value == 6) return PM_ERR_INVALID;
      c->gather_mode = value;
      return PM_OK;
    case PM_OPT_BLOCKS_PER_CU:
      if (value < 1 || value > 64) return PM_ERR_INVALID;
      c->blocks_per_cu = value;
      return PM_OK;
    case PM_OPT_CACHE_INFER_CHUNK:
      if (value < 14 || value > 20) return PM_ERR_INVALID;
      c->cache.infer_chunk_log2 = value;
      return PM_OK;
    case PM_OPT_CACHE_FUSED_INFER:
      if (value < 0 || value > 4) return PM_ERR_INVALID;
      c->cache.fused_infer = value;
      return PM_OK;
    case PM_OPT_TILE_REGIONS:
      if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8) return PM_ERR_INVALID;
      c->tile_regions = value;
      return PM_OK;
    case PM_OPT_SAMPLE_LANES:
      if (value != 0 && value != 1) return PM_ERR_INVALID;
      c->sample_lanes = value;
      return PM_OK;
    case PM_OPT_RENDER_MEDIA:
      if (value != 0 && value != 1) return PM_ERR_INVALID;
      c->render_media = value;
      return PM_OK;
    case PM_OPT_CACHE_EMA:
      if (value < 0 || value > 999) return PM_ERR_INVALID;
      c->cache.ema_decay = (float)value / 1000.0f;
      c->cache.ema_stale = true;
      return PM_OK;
    case PM_OPT_CACHE_SEEDS:
      if (value != 0 && value != 1) return PM_ERR_INVALID;
      c->cache.use_seeds = value;
      return PM_OK;
    case PM_OPT_CACHE_NEAR_GROUPS:
      if (value < 0 || value > 16) return PM_ERR_INVALID;
      c->cache.near_groups = value;
      return PM_OK;
    case PM_OPT_TOP_LEVELS:
      if (value < 0 || value > kMaxTopLevels) return PM_ERR_INVALID;
      c->top_levels = value;
      return PM_OK;
    case PM_OPT_VOXEL_GRID:
      if (value < 0 || value > 2) return PM_ERR_INVALID;
      c->voxel_grid = value;
      return PM_OK;
    default:
      return PM_ERR_INVALID;
  }
}

int pm_stats_get(pm_context* c, pm_stats* out) {
  if (!c || !out) return PM_ERR_INVALID;
  *out = c->stats;
  return PM_OK;
}

int pm_get_records(pm_context* c, int map, float* pos, float* power, float* dir, uint32_t n) {
  if (!c || !pos || !power || !dir) return PM_ERR_INVALID;
  if (!c->emitted) return PM_ERR_STATE;
  const DevMap& m = map == PM_MAP_VOLUME ? c->mv.m : c->ms.m;
  if (n != m.n) return PM_ERR_INVALID;
  return records_to_host(m, pos, power, dir);
}

int pm_set_records(pm_context* c, int map, const float* pos, const float* power, const float* dir, uint32_t n,
                   uint32_t n_emit) {
  if (!c || n_emit == 0 || (n > 0 && (!pos || !power || !dir))) return PM_ERR_INVALID;
  if (map != PM_MAP_SURFACE && map != PM_MAP_VOLUME) return PM_ERR_INVALID;
  MapBufs& mb = map == PM_MAP_VOLUME ? c->mv : c->ms;
  const int bf16 = c->cfg.storage == PM_STORE_BF16 ? 1 : 0;
  int r = alloc_map(mb, n, bf16, c->cfg.leaf_size);
  if (r != PM_OK) return r;
  mb.m.n_emit = n_emit;
  if (n > 0) {
    std::vector<float> x(n), y(n), z(n);
    for (uint32_t i = 0; i < n; ++i) {
      x[i] = pos[3 * i];
      y[i] = pos[3 * i + 1];
      z[i] = pos[3 * i + 2];
    }
    HIPCHK(hipMemcpy(mb.m.x, x.data(), 4ull * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(mb.m.y, y.data(), 4ull * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(mb.m.z, z.data(), 4ull * n, hipMemcpyHostToDevice));
    if (bf16) {
      // 8 BF16 per record: power, direction, 2 zero pads (round to nearest even)
      std::vector<uint16_t> p(8ull * n, 0);
      for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
          p[8ull * i + k] = f32_to_bf16(power[3 * i + k]);
          p[8ull * i + 3 + k] = f32_to_bf16(dir[3 * i + k]);
        }
      HIPCHK(hipMemcpy(mb.m.payload, p.data(), 16ull * n, hipMemcpyHostToDevice));
    } else {
      // 8 FP32 per record: power, direction, 2 zero pads
      std::vector<float> p(8ull * n, 0.0f);
      for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
          p[8ull * i + k] = power[3 * i + k];
          p[8ull * i + 3 + k] = dir[3 * i + k];
        }
      HIPCHK(hipMemcpy(mb.m.payload, p.data(), 32ull * n, hipMemcpyHostToDevice));
    }
  }
  if (map == PM_MAP_VOLUME) {
    c->stats.n_volume = n;
    c->stats.n_emit_volume = n_emit;
  } else {
    c->stats.n_surface = n;
    c->stats.n_emit_surface = n_emit;
  }
  c->emitted = true;
  c->built = false;
  c->vox_ready = false;
  return PM_OK;
}

int pm_get_voxel_table(pm_context* c, int map, uint64_t* table, size_t n) {
  if (!c || !table) return PM_ERR_INVALID;
  if (n != 3ull * kSat * kSat * kSat) return PM_ERR_INVALID;
  if (!c->vox_ready) return PM_ERR_STATE;
  const DevBuf& t = map == PM_MAP_VOLUME ? c->vox_sat_v : c->vox_sat_s;
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(table, t.p, 8ull * n, hipMemcpyDeviceToHost));
  return PM_OK;
}

int pm_get_tree(pm_context* c, int map, uint32_t* perm, float* split, uint32_t* axis, uint32_t n, uint32_t ni) {
  if (!c || !perm) return PM_ERR_INVALID;
  if (!c->built) return PM_ERR_STATE;
  const DevMap& m = map == PM_MAP_VOLUME ? c->mv.m : c->ms.m;
  uint32_t want_ni = m.nl > 0 ? m.nl - 1 : 0;
  if (n != m.n || ni != want_ni) return PM_ERR_INVALID;
  return tree_to_host(m, perm, split, axis);
}

int pm_knn(pm_context* c, int map, const float* q, uint32_t nq, int k, float rcap, uint32_t* idx, float* d2,
           uint32_t* count) {
  if (!c || (nq && (!q || !idx || !d2))) return PM_ERR_INVALID;
  if (k < 1 || k > 64 || !(rcap > 0.0f)) return PM_ERR_INVALID;
  if (!c->built) return PM_ERR_STATE;
  const DevMap& m = map == PM_MAP_VOLUME ? c->mv.m : c->ms.m;
  GatherMap g = gather_map(c, m, k, rcap);
  HIPCHK(launch_knn(g, q, nq, idx, d2, count, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return PM_OK;
}

int pm_device_alloc(void** ptr, size_t bytes) {
  if (!ptr) return PM_ERR_INVALID;
  HIPCHK(hipMalloc(ptr, bytes ? bytes : 1));
  return PM_OK;
}

int pm_device_free(void* ptr) {
  HIPCHK(hipFree(ptr));
  return PM_OK;
}

int pm_memcpy_htod(void* dst, const void* src, size_t bytes) {
  if ((!dst || !src) && bytes) return PM_ERR_INVALID;
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return PM_OK;
}

int pm_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  if ((!dst || !src) && bytes) return PM_ERR_INVALID;
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return PM_OK;
}

int pm_device_count(int* n) {
  if (!n) return PM_ERR_INVALID;
  HIPCHK(hipGetDeviceCount(n));
  return PM_OK;
}

int pm_copy_kernel(void* dst, const void* src, size_t bytes) {
  if ((!dst || !src) && bytes) return PM_ERR_INVALID;
  if (bytes % 16) return PM_ERR_INVALID;
  HIPCHK(launch_copy16(src, dst, bytes, 0));
  HIPCHK(hipDeviceSynchronize());
  return PM_OK;
}

// ---------------------------------------------------------------- legacy ABI
void launch_init_random_numbers_kernel(void) {
  if (g_legacy) {
    pm_destroy(g_legacy);
    g_legacy = nullptr;
  }
  pm_config cfg;
  preset_config(1, &cfg);
  int r = pm_create(&cfg, &g_legacy);
  if (r != PM_OK) legacy_fail("launch_init_random_numbers_kernel", r);
}

pm_context* pm_legacy_context(void) {
  if (!g_legacy) {
    // as launch_init_random_numbers_kernel, but a failure returns NULL instead of exiting
    pm_config cfg;
    preset_config(1, &cfg);
    if (pm_create(&cfg, &g_legacy) != PM_OK) g_legacy = nullptr;
  }
  return g_legacy;
}

static void legacy_configure(float animTime, bool interpolateFlag, bool participatingMediaFlag) {
  if (!g_legacy) launch_init_random_numbers_kernel();
  pm_config& c = g_legacy->cfg;
  bool scene_changed = c.anim_time != animTime || c.media_on != (participatingMediaFlag ? 1 : 0);
  c.anim_time = animTime;
  c.media_on = participatingMediaFlag ? 1 : 0;
  c.estimator = interpolateFlag ? PM_EST_CONE : PM_EST_BOX;
  if (g_legacy->voxel_grid) g_legacy->voxel_grid = interpolateFlag ? 2 : 1;
  if (scene_changed) {
    // the spheres move (or the medium switches) for the eye rays at once; the
    // photon maps stay those of the last emission until the next emit call,
    // as the reference's grid does while its emitFlag is off
    // (callbacksPBO.cpp:60-61, photonMappingKernel.cu:1523-1566)
    int r = setup_scene(g_legacy);
    if (r != PM_OK) legacy_fail("scene setup", r);
  }
}

void launch_emit_photons_kernel(pm_uchar4* pos, unsigned int image_width, unsigned int image_height, float animTime,
                                bool interpolateFlag, bool participatingMediaFlag) {
  (void)pos; (void)image_width; (void)image_height;
  legacy_configure(animTime, interpolateFlag, participatingMediaFlag);
  int r = pm_emit(g_legacy);
  if (r != PM_OK) legacy_fail("launch_emit_photons_kernel (emit)", r);
  r = pm_build(g_legacy);
  if (r != PM_OK) legacy_fail("launch_emit_photons_kernel (build)", r);
}

void launch_photon_mapping_kernel(pm_uchar4* pos, unsigned int image_width, unsigned int image_height,
                                  float animTime, bool interpolateFlag, bool participatingMediaFlag) {
  legacy_configure(animTime, interpolateFlag, participatingMediaFlag);
  if (!pos) legacy_fail("launch_photon_mapping_kernel", PM_ERR_INVALID);
  pm_context* c = g_legacy;
  if (!c->emitted || !c->built) {  // no emission yet (or its trees are missing): emit now
    int r = c->emitted ? PM_OK : pm_emit(c);
    if (r == PM_OK) r = pm_build(c);
    if (r != PM_OK) legacy_fail("launch_photon_mapping_kernel (emit)", r);
  }
  int r = pm_set_render(c, (int)image_width, (int)image_height, 1, c->cfg.k_surface, c->cfg.k_volume,
                        c->cfg.rcap_surface, c->cfg.rcap_volume);
  if (r == PM_OK) r = pm_render(c, nullptr, pos);
  if (r != PM_OK) legacy_fail("launch_photon_mapping_kernel", r);
}

}  // extern "C"
