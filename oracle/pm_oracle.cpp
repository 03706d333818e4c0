// pm_oracle.cpp -- CPU oracle of the photon-mapper hot path.  TEST
// INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker; the product (libpm.so) never
// links or calls it.
//
// What is shared with the GPU path and what is restated here:
//  * shared (per the method definition "the same per-thread logic compiled
//    for the host cores on the same RNG stream"): the per-thread headers in
//    cuda-photon-mapper_amd/csrc/core -- Philox, own sin/cos/log/exp,
//    intersection, sampling, medium, per-photon trace_photon and per-sample
//    render_sample.  Those are pinned independently of the GPU by
//    tests/test_oracle_*.py (Philox known-answer vectors cross-checked against
//    rocRAND's engine, math accuracy against libm in FP64, closed forms,
//    sampling moments, flux conservation).
//  * restated here, independently of the GPU kernels: the record collection
//    and ordering (photon id, event), the photon-cap policy, the canonical
//    left-balanced kd-tree (recursive nth_element), kNN by brute force and by
//    tree descent with a max-heap, the radiance estimators, the pixel loop
//    and an FP64 render.
//
// Reference behaviour followed (photonMappingKernel.cu): emitPhotons
// :1215-1375 (emit -> store -> bounce), storePhoton :1164-1183 (records per
// diffuse hit), gatherPhotons/integrate :314-389,586-600 (surface radiance from
// stored photons), integrateVolumePhotons :831-870 and the march in
// computePixelColor :937-966 (volume in-scatter), frame write :1451-1459.
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <utility>
#include <vector>

#include "../cuda-photon-mapper_amd/csrc/core/pm_kdtree.h"
#include "../cuda-photon-mapper_amd/csrc/core/pm_presets.h"

using namespace pm;

namespace {

struct Map {
  std::vector<float> x, y, z, pr, pg, pb, dx, dy, dz;
  uint32_t n_emit = 0;
  // kd-tree
  uint32_t L = 32, nl = 0;
  std::vector<uint32_t> perm;
  std::vector<float> split;
  std::vector<uint32_t> axis, lcount;
  bool built = false;
  size_t size() const { return x.size(); }
  void push(V3f p, V3f pw, V3f d) {
    x.push_back(p.x); y.push_back(p.y); z.push_back(p.z);
    pr.push_back(pw.x); pg.push_back(pw.y); pb.push_back(pw.z);
    dx.push_back(d.x); dy.push_back(d.y); dz.push_back(d.z);
  }
  float coord(uint32_t i, int a) const { return a == 0 ? x[i] : (a == 1 ? y[i] : z[i]); }
};

struct Ctx {
  pm_config cfg;
  std::vector<Sphere> spheres;
  std::vector<float> density;
  Scene scene;
  TraceParams tp;
  RenderParams rp;
  Map surf, vol;
  Tally9 tally;
  uint32_t n_emit_total = 0;
  int nthreads = 1;
};

// Collects one photon's records in event order (at most 64 events per photon).
struct ArraySink {
  float s[64 * 9], v[64 * 9];
  int ns = 0, nv = 0;
  bool keep_s = true, keep_v = true;  // false: count only (the map is already full)
  void surface(V3f p, V3f pw, V3f d) { put(s, ns, keep_s, p, pw, d); }
  void volume(V3f p, V3f pw, V3f d) { put(v, nv, keep_v, p, pw, d); }
  static void put(float* a, int& n, bool keep, V3f p, V3f pw, V3f d) {
    if (keep) {
      float* r = a + 9 * n;
      r[0] = p.x; r[1] = p.y; r[2] = p.z; r[3] = pw.x; r[4] = pw.y; r[5] = pw.z; r[6] = d.x; r[7] = d.y; r[8] = d.z;
    }
    ++n;
  }
};

struct alignas(64) ThreadOut {
  std::vector<float> s, v;
};

struct PhotonOut {
  std::vector<ThreadOut> per;  // one per thread, in photon-range order
  std::vector<uint32_t> s_cnt, v_cnt;
  std::vector<Tally9> tal;
};

// Trace photons [p0, p1) with `nt` threads; thread t owns a contiguous photon
// range, so concatenating the per-thread outputs gives photon order.
void trace_range(const Ctx& c, uint32_t p0, uint32_t p1, int nt, bool keep_s, bool keep_v, PhotonOut& out) {
  const uint32_t n = p1 - p0;
  out.s_cnt.assign(n, 0);
  out.v_cnt.assign(n, 0);
  out.tal.assign(n, Tally9());
  out.per.assign((size_t)nt, ThreadOut());
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) {
    th.emplace_back([&, t]() {
      uint32_t a = p0 + (uint32_t)((uint64_t)n * t / nt), b = p0 + (uint32_t)((uint64_t)n * (t + 1) / nt);
      std::vector<float> s, v;
      ArraySink* sk = new ArraySink();
      sk->keep_s = keep_s;
      sk->keep_v = keep_v;
      for (uint32_t pid = a; pid < b; ++pid) {
        sk->ns = sk->nv = 0;
        Tally9 tl;
        tally_zero(tl);
        trace_photon(c.scene, c.tp, pid, *sk, tl);
        out.s_cnt[pid - p0] = (uint32_t)sk->ns;
        out.v_cnt[pid - p0] = (uint32_t)sk->nv;
        out.tal[pid - p0] = tl;
        if (keep_s) s.insert(s.end(), sk->s, sk->s + 9 * sk->ns);
        if (keep_v) v.insert(v.end(), sk->v, sk->v + 9 * sk->nv);
      }
      delete sk;
      out.per[(size_t)t].s.swap(s);
      out.per[(size_t)t].v.swap(v);
    });
  }
  for (auto& x : th) x.join();
}

// Record j (in photon order) of a PhotonOut map, walking the per-thread chunks.
struct RecCursor {
  const PhotonOut* po;
  bool vol;
  size_t chunk = 0, off = 0;
  const float* next() {
    for (;;) {
      const std::vector<float>& a = vol ? po->per[chunk].v : po->per[chunk].s;
      if (off < a.size()) {
        const float* r = &a[off];
        off += 9;
        return r;
      }
      ++chunk;
      off = 0;
    }
  }
};

// Appends `take` records from the cursor and skips `skip` more.
void append_records(Map& m, RecCursor& cur, uint64_t take, uint64_t skip) {
  for (uint64_t i = 0; i < take; ++i) {
    const float* q = cur.next();
    m.push(v3(q[0], q[1], q[2]), v3(q[3], q[4], q[5]), v3(q[6], q[7], q[8]));
  }
  for (uint64_t i = 0; i < skip; ++i) cur.next();
}

// Emission.  n_emit > 0: photons [0, n_emit), every record kept (up to the
// caps when set).  n_emit == 0: photons are traced in id order until each map
// holds its cap; a map's normalisation count is (id of the photon that
// produced its cap-th record) + 1, and records beyond the cap are dropped.
int emit(Ctx& c) {
  c.surf = Map();
  c.vol = Map();
  tally_zero(c.tally);
  const bool media = c.tp.media_on != 0;
  const uint64_t cap_s = c.cfg.cap_surface ? c.cfg.cap_surface : UINT64_MAX;
  const uint64_t cap_v = (c.cfg.cap_volume && media) ? c.cfg.cap_volume : UINT64_MAX;
  const bool fixed = c.cfg.n_emit > 0;
  uint32_t pid = 0;
  bool done_s = false, done_v = !media;
  uint32_t chunk = fixed ? c.cfg.n_emit : 65536u;
  // the library's bound: no 32-bit record total may wrap
  const uint64_t max_photons = std::min<uint64_t>(1ull << 31, 0xffffffffull / (uint64_t)c.cfg.max_bounces);
  if (fixed && c.cfg.n_emit > max_photons) return PM_ERR_CAPACITY;
  while (true) {
    uint32_t p1 = fixed ? c.cfg.n_emit : (uint32_t)std::min<uint64_t>((uint64_t)pid + chunk, max_photons);
    if (p1 <= pid) break;
    PhotonOut po;
    trace_range(c, pid, p1, c.nthreads, !done_s, !done_v, po);
    RecCursor cs{&po, false}, cv{&po, true};
    for (uint32_t i = 0; i < p1 - pid; ++i) {
      uint32_t ns = po.s_cnt[i], nv = po.v_cnt[i];
      uint32_t id = pid + i;
      if (!done_s) {
        uint64_t room = cap_s - c.surf.size();
        uint64_t take = std::min<uint64_t>(room, ns);
        append_records(c.surf, cs, take, ns - take);
        if (!fixed && c.surf.size() >= cap_s) {
          done_s = true;
          c.surf.n_emit = id + 1;
        }
      }
      if (!done_v) {
        uint64_t room = cap_v - c.vol.size();
        uint64_t take = std::min<uint64_t>(room, nv);
        append_records(c.vol, cv, take, nv - take);
        if (!fixed && c.vol.size() >= cap_v) {
          done_v = true;
          c.vol.n_emit = id + 1;
        }
      }
      for (int k = 0; k < 9; ++k) c.tally.v[k] += po.tal[i].v[k];
      c.n_emit_total = id + 1;
      if (!fixed && done_s && done_v) break;
    }
    pid = p1;
    if (fixed || (done_s && done_v)) break;
    if (pid >= max_photons) return PM_ERR_CAPACITY;
    chunk = std::min<uint32_t>(chunk * 2u, 1u << 22);
  }
  if (fixed) {
    c.surf.n_emit = c.cfg.n_emit;
    c.vol.n_emit = c.cfg.n_emit;
  }
  if (!media) c.vol.n_emit = c.n_emit_total;
  return 0;
}

// ---------------------------------------------------------------- kd-tree
struct AxisLess {
  const Map* m;
  int a;
  bool operator()(uint32_t i, uint32_t j) const {
    uint32_t ki = float_key(m->coord(i, a)), kj = float_key(m->coord(j, a));
    return ki != kj ? ki < kj : i < j;
  }
};

void build_rec(Map& m, uint32_t h, uint32_t lo, uint32_t hi) {
  if (h >= m.nl - 1) {
    std::sort(m.perm.begin() + lo, m.perm.begin() + hi, AxisLess{&m, 0});
    return;
  }
  float mn[3], mx[3];
  for (int a = 0; a < 3; ++a) {
    mn[a] = m.coord(m.perm[lo], a);
    mx[a] = mn[a];
  }
  for (uint32_t i = lo + 1; i < hi; ++i)
    for (int a = 0; a < 3; ++a) {
      float v = m.coord(m.perm[i], a);
      if (v < mn[a]) mn[a] = v;
      if (v > mx[a]) mx[a] = v;
    }
  int axis = 0;
  float best = mx[0] - mn[0];
  for (int a = 1; a < 3; ++a) {
    float e = mx[a] - mn[a];
    if (e > best) {
      best = e;
      axis = a;
    }
  }
  uint32_t lc = m.L * kd_leaves_under(2u * h + 1u, m.nl);
  uint32_t mid = lo + lc;
  std::nth_element(m.perm.begin() + lo, m.perm.begin() + mid, m.perm.begin() + hi, AxisLess{&m, axis});
  m.split[h] = m.coord(m.perm[mid], axis);
  m.axis[h] = (uint32_t)axis;
  m.lcount[h] = lc;
  build_rec(m, 2u * h + 1u, lo, mid);
  build_rec(m, 2u * h + 2u, mid, hi);
}

void build_tree(Map& m, uint32_t L) {
  const uint32_t n = (uint32_t)m.size();
  m.L = L;
  m.nl = kd_num_leaves(n, L);
  m.perm.resize(n);
  for (uint32_t i = 0; i < n; ++i) m.perm[i] = i;
  uint32_t ni = m.nl > 0 ? m.nl - 1 : 0;
  m.split.assign(ni, 0.0f);
  m.axis.assign(ni, 0);
  m.lcount.assign(ni, 0);
  if (n > 0) build_rec(m, 0, 0, n);
  m.built = true;
}

// ---------------------------------------------------------------- kNN
template <class T>
struct Nb {
  T d2;
  uint32_t idx;
  bool operator<(const Nb& o) const { return d2 != o.d2 ? d2 < o.d2 : idx < o.idx; }
};

template <class T>
T dist2_t(const Map& m, uint32_t i, V3<T> q) {
  T dx = T(m.x[i]) - q.x, dy = T(m.y[i]) - q.y, dz = T(m.z[i]) - q.z;
  return (dx * dx + dy * dy) + dz * dz;
}

template <class T>
struct KnnState {
  std::vector<Nb<T>> heap;  // max-heap under operator<
  int k;
  T rcap2;
  Nb<T> worst() const {
    if ((int)heap.size() < k) return Nb<T>{rcap2, 0xffffffffu};
    return heap.front();
  }
  void offer(T d2, uint32_t idx) {
    if (!(d2 <= rcap2)) return;
    Nb<T> c{d2, idx};
    if (!(c < worst())) return;
    if ((int)heap.size() == k) {
      std::pop_heap(heap.begin(), heap.end());
      heap.pop_back();
    }
    heap.push_back(c);
    std::push_heap(heap.begin(), heap.end());
  }
};

template <class T>
void knn_visit(const Map& m, uint32_t h, uint32_t lo, uint32_t hi, V3<T> q, KnnState<T>& st) {
  if (h >= m.nl - 1) {
    for (uint32_t i = lo; i < hi; ++i) {
      uint32_t p = m.perm[i];
      st.offer(dist2_t(m, p, q), p);
    }
    return;
  }
  int a = (int)m.axis[h];
  T diff = comp(q, a) - T(m.split[h]);
  uint32_t mid = lo + m.lcount[h];
  bool left_first = diff < T(0);
  if (left_first) knn_visit(m, 2 * h + 1, lo, mid, q, st);
  else knn_visit(m, 2 * h + 2, mid, hi, q, st);
  if (diff * diff <= st.worst().d2) {
    if (left_first) knn_visit(m, 2 * h + 2, mid, hi, q, st);
    else knn_visit(m, 2 * h + 1, lo, mid, q, st);
  }
}

template <class T>
std::vector<Nb<T>> knn_tree(const Map& m, V3<T> q, int k, T rcap2) {
  KnnState<T> st;
  st.k = k;
  st.rcap2 = rcap2;
  st.heap.reserve(k);
  if (m.size() > 0) knn_visit(m, 0, 0, (uint32_t)m.size(), q, st);
  std::sort(st.heap.begin(), st.heap.end());
  return st.heap;
}

template <class T>
std::vector<Nb<T>> knn_brute(const Map& m, V3<T> q, int k, T rcap2) {
  std::vector<Nb<T>> all;
  for (uint32_t i = 0; i < (uint32_t)m.size(); ++i) {
    T d2 = dist2_t(m, i, q);
    if (d2 <= rcap2) all.push_back(Nb<T>{d2, i});
  }
  size_t kk = std::min<size_t>((size_t)k, all.size());
  std::partial_sort(all.begin(), all.begin() + kk, all.end());
  all.resize(kk);
  return all;
}

// ---------------------------------------------------------------- estimators
template <class T>
struct OracleGather {
  const Ctx* c;
  bool bf16;  // read power/direction through BF16 rounding (precision variant)
  uint64_t nsurf = 0, nvol = 0;
  T rd(float v) const { return bf16 ? T(bf16_to_f32(f32_to_bf16(v))) : T(v); }

  V3<T> surface(V3<T> x, V3<T> n, V3<T> albedo, V3<T> = V3<T>(), V3<T> = V3<T>()) {
    ++nsurf;
    const Map& m = c->surf;
    const int k = c->cfg.k_surface;
    T rcap = T(c->cfg.rcap_surface);
    T rcap2 = rcap * rcap;
    std::vector<Nb<T>> nb = knn_tree(m, x, k, rcap2);
    if (nb.empty()) return v3<T>(T(0), T(0), T(0));
    T r2 = (int)nb.size() == k ? nb.back().d2 : rcap2;
    const bool cone = c->cfg.estimator == PM_EST_CONE;
    T r = sqrt_(r2);
    NSum3<T> s;
    nsum3_init(s);
    for (const Nb<T>& e : nb) {
      V3<T> dir = v3<T>(rd(m.dx[e.idx]), rd(m.dy[e.idx]), rd(m.dz[e.idx]));
      if (!(dot(dir, n) < T(0))) continue;
      T w = T(1);
      if (cone) w = cone_weight(e.d2, r);
      nsum3_add(s, v3<T>(rd(m.pr[e.idx]), rd(m.pg[e.idx]), rd(m.pb[e.idx])) * w);
    }
    V3<T> sum = nsum3_get(s);
    T norm = surface_norm(r2, cone);
    V3<T> sc = v3<T>(T(c->scene.light_power[0]) / T(m.n_emit), T(c->scene.light_power[1]) / T(m.n_emit),
                     T(c->scene.light_power[2]) / T(m.n_emit));
    return mul(mul(sum, albedo), sc) * norm;
  }

  V3<T> volume(V3<T> x, V3<T> wo, V3<T> = V3<T>()) {
    ++nvol;
    const Map& m = c->vol;
    const int k = c->cfg.k_volume;
    T rcap = T(c->cfg.rcap_volume);
    T rcap2 = rcap * rcap;
    std::vector<Nb<T>> nb = knn_tree(m, x, k, rcap2);
    if (nb.empty()) return v3<T>(T(0), T(0), T(0));
    T r2 = (int)nb.size() == k ? nb.back().d2 : rcap2;
    const T g = T(c->scene.med.g);
    NSum3<T> s;
    nsum3_init(s);
    for (const Nb<T>& e : nb) {
      V3<T> dir = v3<T>(rd(m.dx[e.idx]), rd(m.dy[e.idx]), rd(m.dz[e.idx]));
      T ph = phase_hg(dot(dir, wo), g);
      nsum3_add(s, v3<T>(rd(m.pr[e.idx]), rd(m.pg[e.idx]), rd(m.pb[e.idx])) * ph);
    }
    V3<T> sum = nsum3_get(s);
    T norm = volume_norm(r2);
    V3<T> sc = v3<T>(T(c->scene.light_power[0]) / T(m.n_emit), T(c->scene.light_power[1]) / T(m.n_emit),
                     T(c->scene.light_power[2]) / T(m.n_emit));
    return mul(sum, sc) * norm;
  }
};

// A gather with a constant answer: volume queries return Sv, surface queries
// Sv_s.  With it the march's quadrature weights (transmittance times step
// length) and the segment bounds can be checked against closed forms.
template <class T>
struct ConstGather {
  V3<T> sv, ss;
  uint64_t nsurf = 0, nvol = 0;
  V3<T> surface(V3<T>, V3<T>, V3<T>, V3<T> = V3<T>(), V3<T> = V3<T>()) {
    ++nsurf;
    return ss;
  }
  V3<T> volume(V3<T>, V3<T>, V3<T> = V3<T>()) {
    ++nvol;
    return sv;
  }
};

template <class T>
void render_crop(Ctx& c, int x0, int y0, int w, int h, bool bf16, double* rgb, uint64_t* nq) {
  std::atomic<int> next(0);
  std::vector<std::thread> th;
  std::vector<uint64_t> qs((size_t)c.nthreads * 2, 0);
  for (int t = 0; t < c.nthreads; ++t) {
    th.emplace_back([&, t]() {
      OracleGather<T> g;
      g.c = &c;
      g.bf16 = bf16;
      const int total = w * h, block = 8;
      int b0;
      while ((b0 = next.fetch_add(block)) < total) {
        for (int pi = b0; pi < std::min(total, b0 + block); ++pi) {
          int row = pi / w, i = pi % w;
          uint32_t px = (uint32_t)(x0 + i), py = (uint32_t)(y0 + row);
          NSum3<T> acc;
          nsum3_init(acc);
          for (int s = 0; s < c.rp.spp; ++s) nsum3_add(acc, render_sample<T>(c.scene, c.rp, px, py, (uint32_t)s, g));
          V3<T> v = nsum3_get(acc) * (T(1) / T(c.rp.spp));
          double* o = rgb + ((size_t)row * w + i) * 3;
          o[0] = (double)v.x;
          o[1] = (double)v.y;
          o[2] = (double)v.z;
        }
      }
      qs[(size_t)t * 2] = g.nsurf;
      qs[(size_t)t * 2 + 1] = g.nvol;
    });
  }
  for (auto& x : th) x.join();
  if (nq) {
    nq[0] = nq[1] = 0;
    for (int t = 0; t < c.nthreads; ++t) {
      nq[0] += qs[(size_t)t * 2];
      nq[1] += qs[(size_t)t * 2 + 1];
    }
  }
}

Map& map_of(Ctx& c, int map) { return map == PM_MAP_VOLUME ? c.vol : c.surf; }

// ---------------------------------------------------------------- voxel grids
// The reference's voxel-grid estimate (photonMappingKernel.cu:260-267 voxel
// of a point, :1147-1183 photon into its voxel, :314-389 the surface window,
// :442-583 the bilinear blend, :831-870 the volume window), restated with the
// library's units: record power in 2^-24 fixed point, 64-bit sums, the room
// is the unit cube.  Window sums are taken cell by cell from the grid, not
// from summed-volume tables.
constexpr int OV = 32;
constexpr double OFIX = 16777216.0;

int ovox(float p) {
  const int v = (int)(p * 32.0f);
  return std::min(std::max(v, 0), OV - 1);
}

struct VoxGrid {
  std::vector<uint64_t> g;  // [z][y][x][channel]
  uint64_t at(int x, int y, int z, int ch) const { return g[(((size_t)z * OV + y) * OV + x) * 3 + ch]; }
};

VoxGrid voxel_grid(const Map& m, bool bf16) {
  VoxGrid G;
  G.g.assign((size_t)OV * OV * OV * 3, 0);
  auto rd = [bf16](float v) { return bf16 ? bf16_to_f32(f32_to_bf16(v)) : v; };
  for (size_t i = 0; i < m.size(); ++i) {
    const size_t cell = (((size_t)ovox(m.z[i]) * OV + ovox(m.y[i])) * OV + ovox(m.x[i])) * 3;
    const float p[3] = {rd(m.pr[i]), rd(m.pg[i]), rd(m.pb[i])};
    for (int k = 0; k < 3; ++k) G.g[cell + k] += (uint64_t)((double)p[k] * OFIX + 0.5);
  }
  return G;
}

// Summed-volume table by the 3-D recurrence (inclusion-exclusion of the 7
// neighbours), modulo 2^64.
std::vector<uint64_t> voxel_table(const VoxGrid& G) {
  const int S = OV + 1;
  std::vector<uint64_t> t((size_t)S * S * S * 3, 0);
  auto T = [&](int x, int y, int z, int ch) -> uint64_t& { return t[(((size_t)z * S + y) * S + x) * 3 + ch]; };
  for (int z = 1; z < S; ++z)
    for (int y = 1; y < S; ++y)
      for (int x = 1; x < S; ++x)
        for (int ch = 0; ch < 3; ++ch)
          T(x, y, z, ch) = G.at(x - 1, y - 1, z - 1, ch) + T(x - 1, y, z, ch) + T(x, y - 1, z, ch) +
                           T(x, y, z - 1, ch) - T(x - 1, y - 1, z, ch) - T(x - 1, y, z - 1, ch) -
                           T(x, y - 1, z - 1, ch) + T(x - 1, y - 1, z - 1, ch);
  return t;
}

template <class T>
struct VoxelOracleGather {
  const VoxGrid* gs;
  const VoxGrid* gv;
  double ss[3], sv[3];  // light power / photons emitted (FP32 quotient, as the library forms it)
  bool interp = false;
  uint64_t nsurf = 0, nvol = 0;

  // Flux density of the window around in-plane cell (cb, cc) on the wall
  // layer `layer` across axis a: the reference's [v - 3, v + 3) clamped to
  // [0, 32] in each in-plane axis.
  void window(int a, int layer, int cb, int cc, double e[3]) const {
    const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
    const int blo = std::max(cb - 3, 0), bhi = std::min(cb + 3, OV);
    const int clo = std::max(cc - 3, 0), chi = std::min(cc + 3, OV);
    uint64_t q[3] = {0, 0, 0};
    for (int u = blo; u < bhi; ++u)
      for (int v = clo; v < chi; ++v) {
        int ix[3];
        ix[a] = layer; ix[b] = u; ix[c] = v;
        for (int ch = 0; ch < 3; ++ch) q[ch] += gs->at(ix[0], ix[1], ix[2], ch);
      }
    const double area = (double)((bhi - blo) * (chi - clo)) / (double)(OV * OV);
    for (int ch = 0; ch < 3; ++ch) e[ch] = (double)q[ch] / OFIX * ss[ch] / area;
  }

  V3<T> surface(V3<T> x, V3<T> n, V3<T> albedo, V3<T> = V3<T>(), V3<T> = V3<T>()) {
    ++nsurf;
    // walls only: a normal along one axis
    const double nn[3] = {std::fabs((double)n.x), std::fabs((double)n.y), std::fabs((double)n.z)};
    int a = -1;
    for (int k = 0; k < 3; ++k)
      if (nn[k] == 1.0 && nn[(k + 1) % 3] == 0.0 && nn[(k + 2) % 3] == 0.0) a = k;
    if (a < 0) return v3<T>(T(0), T(0), T(0));  // spheres: no estimate
    const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
    const double xs[3] = {(double)x.x, (double)x.y, (double)x.z};
    const int layer = xs[a] > 0.5 ? OV - 1 : 0;
    double e[3];
    if (!interp) {
      window(a, layer, ovox((float)xs[b]), ovox((float)xs[c]), e);
    } else {
      // blend the windows at the 4 voxel centres (i + 1/2) / 32 around x
      const double ub = xs[b] * OV - 0.5, uc = xs[c] * OV - 0.5;
      const double fb0 = ub < 0.0 ? -1.0 : std::trunc(ub), fc0 = uc < 0.0 ? -1.0 : std::trunc(uc);
      const double fb = std::min(std::max(ub - fb0, 0.0), 1.0), fc = std::min(std::max(uc - fc0, 0.0), 1.0);
      const int ib = (int)fb0, ic = (int)fc0;
      auto cl = [](int v) { return std::min(std::max(v, 0), OV - 1); };
      double w00[3], w10[3], w01[3], w11[3];
      window(a, layer, cl(ib), cl(ic), w00);
      window(a, layer, cl(ib + 1), cl(ic), w10);
      window(a, layer, cl(ib), cl(ic + 1), w01);
      window(a, layer, cl(ib + 1), cl(ic + 1), w11);
      for (int ch = 0; ch < 3; ++ch) {
        const double lo = w00[ch] + fb * (w10[ch] - w00[ch]);
        const double hi = w01[ch] + fb * (w11[ch] - w01[ch]);
        e[ch] = lo + fc * (hi - lo);
      }
    }
    const double ip = 0.318309886183790671537767526745;  // 1 / pi
    return v3<T>(T((double)albedo.x * ip * e[0]), T((double)albedo.y * ip * e[1]), T((double)albedo.z * ip * e[2]));
  }

  // The reference's interior window [v - 3, v + 3) clamped to [1, 31) per axis.
  V3<T> volume(V3<T> x, V3<T>, V3<T> = V3<T>()) {
    ++nvol;
    int lo[3], hi[3];
    const float xs[3] = {(float)x.x, (float)x.y, (float)x.z};
    for (int a = 0; a < 3; ++a) {
      const int v = ovox(xs[a]);
      lo[a] = std::min(std::max(v - 3, 1), OV - 1);
      hi[a] = std::min(std::max(v + 3, 1), OV - 1);
    }
    const int cells = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
    if (cells <= 0) return v3<T>(T(0), T(0), T(0));
    uint64_t q[3] = {0, 0, 0};
    for (int zz = lo[2]; zz < hi[2]; ++zz)
      for (int yy = lo[1]; yy < hi[1]; ++yy)
        for (int xx = lo[0]; xx < hi[0]; ++xx)
          for (int ch = 0; ch < 3; ++ch) q[ch] += gv->at(xx, yy, zz, ch);
    const double vol = (double)cells / (double)(OV * OV * OV);
    const double i4p = 0.0795774715459476678844418816863;  // 1 / (4 pi)
    double s[3];
    for (int ch = 0; ch < 3; ++ch) s[ch] = (double)q[ch] / OFIX * sv[ch] / vol * i4p;
    return v3<T>(T(s[0]), T(s[1]), T(s[2]));
  }
};

template <class T>
VoxelOracleGather<T> voxel_gather(const Ctx& c, const VoxGrid& gs, const VoxGrid& gv, bool interp) {
  VoxelOracleGather<T> g;
  g.gs = &gs;
  g.gv = &gv;
  g.interp = interp;
  for (int k = 0; k < 3; ++k) {
    g.ss[k] = (double)(c.cfg.light_power[k] / (float)(c.surf.n_emit ? c.surf.n_emit : 1u));
    g.sv[k] = (double)(c.cfg.light_power[k] / (float)(c.vol.n_emit ? c.vol.n_emit : 1u));
  }
  return g;
}

}  // namespace

extern "C" {

int orc_config_preset(int preset, pm_config* out) {
  if (!out) return PM_ERR_INVALID;
  return preset_config(preset, out) == 0 ? 0 : PM_ERR_INVALID;
}

void* orc_create(const pm_config* cfg, int nthreads) {
  if (!cfg || validate_config(*cfg) != 0) return nullptr;
  Ctx* c = new Ctx();
  c->cfg = *cfg;
  c->spheres = build_spheres(*cfg);
  c->density = build_density(*cfg);
  c->scene = make_scene(*cfg, c->spheres.data(), c->density.empty() ? nullptr : c->density.data());
  c->tp = trace_params(*cfg);
  c->rp = render_params(*cfg);
  c->nthreads = nthreads > 0 ? nthreads : 1;
  return c;
}

void orc_destroy(void* h) { delete (Ctx*)h; }

int orc_set_render(void* h, int width, int height, int spp, int k_s, int k_v, float rcap_s, float rcap_v) {
  Ctx* c = (Ctx*)h;
  c->cfg.width = width; c->cfg.height = height; c->cfg.spp = spp;
  c->cfg.k_surface = k_s; c->cfg.k_volume = k_v;
  c->cfg.rcap_surface = rcap_s; c->cfg.rcap_volume = rcap_v;
  if (validate_config(c->cfg) != 0) return PM_ERR_INVALID;
  c->rp = render_params(c->cfg);
  return 0;
}

int orc_emit(void* h) { return emit(*(Ctx*)h); }

// counts[0..4] = n_surface, n_volume, n_emit_surface, n_emit_volume, n_emit_total
int orc_counts(void* h, uint32_t* counts) {
  Ctx* c = (Ctx*)h;
  counts[0] = (uint32_t)c->surf.size();
  counts[1] = (uint32_t)c->vol.size();
  counts[2] = c->surf.n_emit;
  counts[3] = c->vol.n_emit;
  counts[4] = c->n_emit_total;
  return 0;
}

// tally[0..11] = emitted, absorbed, escaped, rr_gain (RGB each), watts.
int orc_tally(void* h, double* out) {
  Ctx* c = (Ctx*)h;
  for (int k = 0; k < 3; ++k) {
    double w = (double)c->scene.light_power[k] / (double)c->n_emit_total;
    out[k] = (double)c->scene.light_power[k];
    out[3 + k] = c->tally.v[k] * w;
    out[6 + k] = c->tally.v[3 + k] * w;
    out[9 + k] = c->tally.v[6 + k] * w;
  }
  return 0;
}

int orc_get_records(void* h, int map, float* pos, float* power, float* dir, uint32_t n) {
  Map& m = map_of(*(Ctx*)h, map);
  if (n != m.size()) return PM_ERR_INVALID;
  for (uint32_t i = 0; i < n; ++i) {
    pos[3 * i] = m.x[i]; pos[3 * i + 1] = m.y[i]; pos[3 * i + 2] = m.z[i];
    power[3 * i] = m.pr[i]; power[3 * i + 1] = m.pg[i]; power[3 * i + 2] = m.pb[i];
    dir[3 * i] = m.dx[i]; dir[3 * i + 1] = m.dy[i]; dir[3 * i + 2] = m.dz[i];
  }
  return 0;
}

// Replace a map's records (tests use hand-placed photon sets).
int orc_set_records(void* h, int map, const float* pos, const float* power, const float* dir, uint32_t n,
                    uint32_t n_emit) {
  Map& m = map_of(*(Ctx*)h, map);
  m = Map();
  for (uint32_t i = 0; i < n; ++i)
    m.push(v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]), v3(power[3 * i], power[3 * i + 1], power[3 * i + 2]),
           v3(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]));
  m.n_emit = n_emit;
  return 0;
}

int orc_build(void* h) {
  Ctx* c = (Ctx*)h;
  build_tree(c->surf, (uint32_t)c->cfg.leaf_size);
  build_tree(c->vol, (uint32_t)c->cfg.leaf_size);
  return 0;
}

int orc_get_tree(void* h, int map, uint32_t* perm, float* split, uint32_t* axis, uint32_t n, uint32_t ni) {
  Map& m = map_of(*(Ctx*)h, map);
  if (!m.built || n != m.size() || ni != (m.nl ? m.nl - 1 : 0)) return PM_ERR_INVALID;
  if (n) memcpy(perm, m.perm.data(), sizeof(uint32_t) * n);
  if (ni) {
    memcpy(split, m.split.data(), sizeof(float) * ni);
    memcpy(axis, m.axis.data(), sizeof(uint32_t) * ni);
  }
  return 0;
}

// kNN for nq host queries (xyz); mode 0 = tree, 1 = brute force.  Outputs
// nq*k (idx, d2) sorted by (d2, idx), padded with 0xffffffff / +inf.
int orc_knn(void* h, int map, const float* q, uint32_t nq, int k, float rcap, int mode, uint32_t* idx, float* d2,
            uint32_t* count) {
  Ctx* c = (Ctx*)h;
  Map& m = map_of(*c, map);
  if (mode == 0 && !m.built) return PM_ERR_STATE;
  const float rcap2 = rcap * rcap;
  std::atomic<uint32_t> next(0);
  std::vector<std::thread> th;
  for (int t = 0; t < c->nthreads; ++t) {
    th.emplace_back([&]() {
      uint32_t i;
      while ((i = next.fetch_add(1)) < nq) {
        V3f qq = v3(q[3 * i], q[3 * i + 1], q[3 * i + 2]);
        std::vector<Nb<float>> nb = mode == 0 ? knn_tree(m, qq, k, rcap2) : knn_brute(m, qq, k, rcap2);
        for (int j = 0; j < k; ++j) {
          bool ok = j < (int)nb.size();
          idx[(size_t)i * k + j] = ok ? nb[j].idx : 0xffffffffu;
          d2[(size_t)i * k + j] = ok ? nb[j].d2 : u2f(0x7f800000u);
        }
        if (count) count[i] = (uint32_t)nb.size();
      }
    });
  }
  for (auto& x : th) x.join();
  return 0;
}

// Radiance of the crop (x0, y0, w, h) of the configured frame; precision 0 =
// FP32 (the parity target), 1 = FP64 (the accuracy reference).  bf16 != 0
// reads photon power and direction through BF16 rounding.  nq (may be NULL)
// receives the surface and volume query counts.
int orc_render(void* h, int x0, int y0, int w, int hh, int precision, int bf16, double* rgb, uint64_t* nq) {
  Ctx* c = (Ctx*)h;
  if (!c->surf.built || !c->vol.built) return PM_ERR_STATE;
  if (precision == 0) render_crop<float>(*c, x0, y0, w, hh, bf16 != 0, rgb, nq);
  else render_crop<double>(*c, x0, y0, w, hh, bf16 != 0, rgb, nq);
  return 0;
}

// The voxel-grid estimate's grids ([32^3][3]) and summed-volume tables
// ([33^3][3]) of both maps, power through BF16 rounding when bf16 (any output
// may be NULL).
int orc_voxel_tables(void* h, int bf16, uint64_t* grid_s, uint64_t* grid_v, uint64_t* sat_s, uint64_t* sat_v) {
  Ctx* c = (Ctx*)h;
  const VoxGrid gs = voxel_grid(c->surf, bf16 != 0), gv = voxel_grid(c->vol, bf16 != 0);
  if (grid_s) memcpy(grid_s, gs.g.data(), gs.g.size() * 8);
  if (grid_v) memcpy(grid_v, gv.g.data(), gv.g.size() * 8);
  if (sat_s) {
    const std::vector<uint64_t> t = voxel_table(gs);
    memcpy(sat_s, t.data(), t.size() * 8);
  }
  if (sat_v) {
    const std::vector<uint64_t> t = voxel_table(gv);
    memcpy(sat_v, t.data(), t.size() * 8);
  }
  return 0;
}

// Voxel-grid estimates at n points in FP64: kind 0 = surface (normal n,
// albedo alb, 3 floats each), 1 = volume.  out: 3 doubles per point.
int orc_voxel_query(void* h, int kind, int interp, int bf16, const float* x, const float* nrm, const float* alb,
                    uint32_t n, double* out) {
  Ctx* c = (Ctx*)h;
  const VoxGrid gs = voxel_grid(c->surf, bf16 != 0), gv = voxel_grid(c->vol, bf16 != 0);
  VoxelOracleGather<double> g = voxel_gather<double>(*c, gs, gv, interp != 0);
  for (uint32_t i = 0; i < n; ++i) {
    const V3<double> p = v3<double>(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    V3<double> r;
    if (kind == 0)
      r = g.surface(p, v3<double>(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]),
                    v3<double>(alb[3 * i], alb[3 * i + 1], alb[3 * i + 2]));
    else
      r = g.volume(p, v3<double>(0.0, 0.0, 1.0));
    out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
  }
  return 0;
}

// The render of the voxel-grid estimate (FP32 frame as the library's), the
// pixel loop of orc_render.  Needs the records only.
int orc_render_voxel(void* h, int x0, int y0, int w, int hh, int interp, int bf16, double* rgb, uint64_t* nq) {
  Ctx* c = (Ctx*)h;
  const VoxGrid gs = voxel_grid(c->surf, bf16 != 0), gv = voxel_grid(c->vol, bf16 != 0);
  std::atomic<int> next(0);
  std::vector<std::thread> th;
  std::vector<uint64_t> qs((size_t)c->nthreads * 2, 0);
  for (int t = 0; t < c->nthreads; ++t) {
    th.emplace_back([&, t]() {
      VoxelOracleGather<float> g = voxel_gather<float>(*c, gs, gv, interp != 0);
      int pi;
      while ((pi = next.fetch_add(1)) < w * hh) {
        const int row = pi / w, i = pi % w;
        NSum3<float> acc;
        nsum3_init(acc);
        for (int s = 0; s < c->rp.spp; ++s)
          nsum3_add(acc, render_sample<float>(c->scene, c->rp, (uint32_t)(x0 + i), (uint32_t)(y0 + row),
                                              (uint32_t)s, g));
        const V3f v = nsum3_get(acc) * (1.0f / (float)c->rp.spp);
        double* o = rgb + ((size_t)row * w + i) * 3;
        o[0] = v.x; o[1] = v.y; o[2] = v.z;
      }
      qs[(size_t)t * 2] = g.nsurf;
      qs[(size_t)t * 2 + 1] = g.nvol;
    });
  }
  for (auto& x : th) x.join();
  if (nq) {
    nq[0] = nq[1] = 0;
    for (int t = 0; t < c->nthreads; ++t) {
      nq[0] += qs[(size_t)t * 2];
      nq[1] += qs[(size_t)t * 2 + 1];
    }
  }
  return 0;
}

// Per-thread primitives exposed for the oracle's own unit tests.
int orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out, uint32_t n) {
  for (uint32_t i = 0; i < n; ++i) {
    U4 r = philox4x32_10(u4(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3]), key[2 * i], key[2 * i + 1]);
    out[4 * i] = r.x; out[4 * i + 1] = r.y; out[4 * i + 2] = r.z; out[4 * i + 3] = r.w;
  }
  return 0;
}

int orc_math(int fn, const float* x, float* y, uint32_t n) {
  for (uint32_t i = 0; i < n; ++i) {
    float s, c;
    switch (fn) {
      case 0: sincos_turns(x[i], &s, &c); y[i] = s; break;
      case 1: sincos_turns(x[i], &s, &c); y[i] = c; break;
      case 2: y[i] = log_(x[i]); break;
      case 3: y[i] = exp_(x[i]); break;
      default: return PM_ERR_INVALID;
    }
  }
  return 0;
}

// Direction samplers: kind 0 sphere, 1 cosine about n=(p0,p1,p2), 2 HG about
// d=(p0,p1,p2) with g=p3.  u holds 2 uniforms per sample.
int orc_sample(int kind, const float* p, const float* u, float* out, uint32_t n) {
  V3f a = v3(p[0], p[1], p[2]);
  for (uint32_t i = 0; i < n; ++i) {
    V3f d;
    if (kind == 0) d = sample_sphere(u[2 * i], u[2 * i + 1]);
    else if (kind == 1) d = sample_cosine(a, u[2 * i], u[2 * i + 1]);
    else d = sample_hg(a, p[3], u[2 * i], u[2 * i + 1]);
    out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
  }
  return 0;
}

// Closest hit for n rays (o, d) against the configured scene: t, kind, id.
int orc_intersect(void* h, const float* o, const float* d, float* t, int* kind, int* id, uint32_t n) {
  Ctx* c = (Ctx*)h;
  for (uint32_t i = 0; i < n; ++i) {
    Hit<float> hh = intersect(c->scene, v3(o[3 * i], o[3 * i + 1], o[3 * i + 2]), v3(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
    t[i] = hh.t; kind[i] = hh.kind; id[i] = hh.id;
  }
  return 0;
}

float orc_fresnel(float cos_i, float eta) {
  float ct;
  return fresnel(cos_i, eta, &ct);
}

float orc_phase_hg(float c, float g) { return phase_hg(c, g); }

int orc_density(void* h, const float* p, float* out, uint32_t n) {
  Ctx* c = (Ctx*)h;
  for (uint32_t i = 0; i < n; ++i) out[i] = density_at(c->scene.med, v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]));
  return 0;
}

int orc_noise_grid(void* h, float* out, uint32_t n) {
  Ctx* c = (Ctx*)h;
  if (n != c->density.size()) return PM_ERR_INVALID;
  if (n) memcpy(out, c->density.data(), sizeof(float) * n);
  return 0;
}

// Free-flight sample for photon pid, event ev from o along d, bounded by tmax.
int orc_collision(void* h, uint32_t pid, int ev, const float* o, const float* d, float tmax, float* tc) {
  Ctx* c = (Ctx*)h;
  float t = -1.0f;
  bool hit = sample_collision(c->scene.med, c->tp.seed, c->tp.max_delta_iters, pid, ev, v3(o[0], o[1], o[2]),
                              v3(d[0], d[1], d[2]), tmax, &t);
  *tc = hit ? t : -1.0f;
  return hit ? 1 : 0;
}

uint32_t orc_kd_leaves_under(uint32_t c, uint32_t nl) { return kd_leaves_under(c, nl); }

// The frame-write rule (tonemap_u8) for n FP32 values.
int orc_tonemap(const float* v, uint32_t n, float exposure, uint8_t* out) {
  for (uint32_t i = 0; i < n; ++i) out[i] = tonemap_u8(v[i], exposure);
  return 0;
}

// Replaces the value-noise density grid (res^3 vertex values, x fastest) of a
// MED_NOISE context: the march and delta tracking then read `grid`.
int orc_set_density(void* h, const float* grid, uint32_t n) {
  Ctx* c = (Ctx*)h;
  if (c->scene.med.type != MED_NOISE || n != c->density.size()) return PM_ERR_INVALID;
  memcpy(c->density.data(), grid, sizeof(float) * n);
  return 0;
}

// march_segment over [t0, t1) along o + t d with the march offset xi, a
// constant in-scatter S and incoming throughput thr: L (the weighted sum) and
// the outgoing throughput; tt64 != 0 accumulates the transmittance in FP64.
int orc_march(void* h, int tt64, const float* o, const float* d, float t0, float t1, float xi, const float* S,
              const float* thr_in, float* L_out, float* thr_out, uint32_t* nq) {
  Ctx* c = (Ctx*)h;
  if (c->scene.med.type == MED_NONE) return PM_ERR_INVALID;
  ConstGather<float> g;
  g.sv = v3(S[0], S[1], S[2]);
  g.ss = v3(0.0f, 0.0f, 0.0f);
  V3f thr = v3(thr_in[0], thr_in[1], thr_in[2]);
  NSum3<float> L;
  nsum3_init(L);
  V3f oo = v3(o[0], o[1], o[2]), dd = v3(d[0], d[1], d[2]);
  if (tt64) march_segment<float, double>(c->scene, c->rp, oo, dd, t0, t1, xi, thr, L, g);
  else march_segment<float, float>(c->scene, c->rp, oo, dd, t0, t1, xi, thr, L, g);
  V3f l = nsum3_get(L);
  L_out[0] = l.x; L_out[1] = l.y; L_out[2] = l.z;
  thr_out[0] = thr.x; thr_out[1] = thr.y; thr_out[2] = thr.z;
  if (nq) *nq = (uint32_t)g.nvol;
  return 0;
}

// One eye sample (pixel px, py, sample s) of the configured frame with a
// constant gather: volume queries return Sv, surface queries Ss (each times
// its path weight).  Counts: nq[0] surface, nq[1] volume queries.
int orc_render_const(void* h, uint32_t px, uint32_t py, uint32_t s, const float* Sv, const float* Ss, float* out,
                     uint32_t* nq) {
  Ctx* c = (Ctx*)h;
  ConstGather<float> g;
  g.sv = v3(Sv[0], Sv[1], Sv[2]);
  g.ss = v3(Ss[0], Ss[1], Ss[2]);
  V3f v = render_sample<float>(c->scene, c->rp, px, py, s, g);
  out[0] = v.x; out[1] = v.y; out[2] = v.z;
  if (nq) {
    nq[0] = (uint32_t)g.nsurf;
    nq[1] = (uint32_t)g.nvol;
  }
  return 0;
}

// Camera ray of pixel (px, py), sample s: origin, direction, march offset.
int orc_camera_ray(void* h, uint32_t px, uint32_t py, uint32_t s, float* o, float* d, float* xi) {
  Ctx* c = (Ctx*)h;
  V3f oo, dd;
  camera_ray(c->rp, px, py, s, &oo, &dd, xi);
  o[0] = oo.x; o[1] = oo.y; o[2] = oo.z;
  d[0] = dd.x; d[1] = dd.y; d[2] = dd.z;
  return 0;
}

// validate_config of the library's headers (0 or PM_ERR_INVALID).
int orc_validate(const pm_config* cfg) { return cfg ? validate_config(*cfg) : PM_ERR_INVALID; }

// One radiance estimate (FP64 when fp64 != 0) at x: which 0 = surface with
// normal v and albedo alb, 1 = volume toward direction v.  out = RGB.
int orc_gather(void* h, int which, const float* x, const float* v, const float* alb, int fp64, double* out) {
  Ctx* c = (Ctx*)h;
  if (!c->surf.built || !c->vol.built) return PM_ERR_STATE;
  if (fp64) {
    OracleGather<double> g;
    g.c = c;
    g.bf16 = false;
    V3d xx = v3((double)x[0], (double)x[1], (double)x[2]), vv = v3((double)v[0], (double)v[1], (double)v[2]);
    V3d r = which == 0 ? g.surface(xx, vv, v3((double)alb[0], (double)alb[1], (double)alb[2])) : g.volume(xx, vv);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
  } else {
    OracleGather<float> g;
    g.c = c;
    g.bf16 = false;
    V3f xx = v3(x[0], x[1], x[2]), vv = v3(v[0], v[1], v[2]);
    V3f r = which == 0 ? g.surface(xx, vv, v3(alb[0], alb[1], alb[2])) : g.volume(xx, vv);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
  }
  return 0;
}

}  // extern "C"
