// pm_lane_knn2.h -- per-lane exact kNN radiance with a 4-byte heap.
//
// The gather of pm_lane_knn.h keeps (d^2, photon index) keys, 8 bytes per
// heap slot; at k = 64 a wave's heaps take 32 KB of LDS, which caps a compute
// unit at 5 waves in flight for a traversal that is bound by dependent
// loads.  This variant keeps only the FP32 bits of d^2 (16 KB per wave at
// k = 64, so twice the waves in flight) and recovers the exact neighbour set
// in a second, fixed-radius pass:
//   pass 1: the multiset of the k smallest d^2 under the search bound; its
//           maximum D is the k-th neighbour's squared distance;
//   pass 2: every point with d^2 < D contributes; of the points with
//           d^2 == D, the (k - #closer) with the smallest photon indices
//           contribute (the (d^2, index) order of the 8-byte gather).
// The neighbour set is therefore identical to pm_lane_knn.h's, and so are
// r_k and the estimate up to the order of the (compensated) sum.
#pragma once

#include "pm_lane_knn.h"

namespace pm {

// Max-heap of FP32 d^2 bits; slot j of this lane at H[j * 64 + lane].
struct LaneHeap32 {
  uint32_t* H;
  int lane;
  int k;
  int size;
  __device__ uint32_t& h(int j) { return H[j * 64 + lane]; }
  __device__ void push(uint32_t v) {
    int j = size++;
    while (j > 0) {
      int p = (j - 1) >> 1;
      uint32_t pv = h(p);
      if (pv >= v) break;
      h(j) = pv;
      j = p;
    }
    h(j) = v;
  }
  __device__ void replace_root(uint32_t v) {
    int j = 0;
    for (;;) {
      int c = 2 * j + 1;
      if (c >= k) break;
      uint32_t cv = h(c);
      if (c + 1 < k) {
        uint32_t dv = h(c + 1);
        if (dv > cv) {
          ++c;
          cv = dv;
        }
      }
      if (cv <= v) break;
      h(j) = cv;
      j = c;
    }
    h(j) = v;
  }
};

// Stackless near-first traversal (the loop structure of lane_knn) that hands
// the leaf points, in batches of up to NB (loads issued together), to a visitor:
//   uint32_t V::bound()  -- a far child is visited iff diff^2 <= u2f(bound())
//   template <int NB> void V::points(const float4* p, const uint32_t* dh, int cnt)
template <int NB = 8, class V>
__device__ __forceinline__ void lane_visit(const GatherMap& g, float qx, float qy, float qz, V& v) {
  const uint32_t n = g.n, nl = g.nl;
  if (n == 0) return;
  uint32_t h = 0, lo = 0, hi = n, from = 0;
  bool down = true, done = false;
  for (;;) {
    for (;;) {
      const bool step = !done && !(down && h + 1u >= nl);
      if (ballot(step) == 0ull) break;
      if (!step) continue;
      if (down) {
        KdNode nd = g.nodes[h];
        int a = kd_axis(nd.meta);
        float qa = a == 0 ? qx : (a == 1 ? qy : qz);
        uint32_t mid = lo + kd_left_count(nd.meta);
        if (qa - nd.split < 0.0f) {
          h = 2u * h + 1u;
          hi = mid;
        } else {
          h = 2u * h + 2u;
          lo = mid;
        }
        continue;
      }
      if (from == 0u) {
        done = true;
        continue;
      }
      uint32_t p = (from - 1u) >> 1;
      KdNode nd = g.nodes[p];
      int a = kd_axis(nd.meta);
      float qa = a == 0 ? qx : (a == 1 ? qy : qz);
      float diff = qa - nd.split;
      uint32_t nearc = diff < 0.0f ? 2u * p + 1u : 2u * p + 2u;
      if (from == nearc && diff * diff <= u2f(v.bound())) {
        uint2 r = g.nrange[p];
        uint32_t mid = r.x + kd_left_count(nd.meta);
        if (nearc == 2u * p + 1u) {
          h = 2u * p + 2u;
          lo = mid;
          hi = r.y;
        } else {
          h = 2u * p + 1u;
          lo = r.x;
          hi = mid;
        }
        down = true;
      } else {
        from = p;
      }
    }
    const bool at_leaf = !done && down && h + 1u >= nl;
    if (ballot(at_leaf) == 0ull) break;
    if (!at_leaf) continue;
    for (uint32_t b0 = lo; b0 < hi; b0 += (uint32_t)NB) {
      float4 p[NB];
      const float4* q = g.tpts + b0;  // padded: entries past hi are read, not used
#pragma unroll
      for (int j = 0; j < NB; ++j) p[j] = q[j];
      uint32_t dh[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) dh[j] = f2u(dist2(qx, qy, qz, p[j].x, p[j].y, p[j].z));
      const uint32_t left = hi - b0;
      v.template points<NB>(p, dh, left < (uint32_t)NB ? (int)left : NB);
    }
    from = h;
    down = false;
  }
}

// Pass 1: the k smallest d^2 (as a multiset) under an inclusive bound.
struct Knn32Visitor {
  LaneHeap32* hp;
  uint32_t wh;  // prune bound: the search bound until the heap is full, then its root
  __device__ uint32_t bound() const { return wh; }
  template <int NB>
  __device__ void points(const float4*, const uint32_t* dh, int cnt) {
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < cnt) {
        const uint32_t d = dh[j];
        if (hp->size < hp->k) {
          if (d <= wh) {
            hp->push(d);
            if (hp->size == hp->k) wh = hp->h(0);
          }
        } else if (d < wh) {
          hp->replace_root(d);
          wh = hp->h(0);
        }
      }
    }
  }
};

// Pass 2: visits every point with d^2 below (or, `inclusive`, at most) D and
// hands it to `add`; with ties tracked, counts the points strictly below D and
// the points at D with the smallest index among them.
template <class Add>
struct SumVisitor {
  uint32_t D;
  bool inclusive;
  uint32_t n_less, n_tie, tie_min;
  Add* add;
  __device__ uint32_t bound() const { return D; }
  template <int NB>
  __device__ void points(const float4* p, const uint32_t* dh, int cnt) {
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < cnt) {
        const uint32_t d = dh[j];
        const uint32_t idx = f2u(p[j].w);
        if (d < D || (inclusive && d == D)) {
          (*add)(idx, d);
          ++n_less;
        } else if (d == D) {
          ++n_tie;
          tie_min = idx < tie_min ? idx : tie_min;
        }
      }
    }
  }
};

// Smallest index above `last` among the points at exactly D (rare path).
struct TieVisitor {
  uint32_t D, last, best;
  __device__ uint32_t bound() const { return D; }
  template <int NB>
  __device__ void points(const float4* p, const uint32_t* dh, int cnt) {
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < cnt && dh[j] == D) {
        const uint32_t idx = f2u(p[j].w);
        if (idx > last && idx < best) best = idx;
      }
    }
  }
};

// Runs pass 1 (with the bounded-then-capped retry of the volume gather when
// seed_h is below cap_h) and pass 2, calling add(idx, d2 bits) once per
// neighbour.  Returns m (neighbours found, <= k) and sets r2_out to the
// estimator's squared radius (d^2 of the k-th neighbour, or the cap).
template <class Add>
__device__ __forceinline__ int knn32_gather(const GatherMap& g, LaneHeap32& hp, float qx, float qy, float qz,
                                            uint32_t seed_h, uint32_t cap_h, Add& add, float& r2_out) {
  hp.k = g.k;
  hp.size = 0;
  Knn32Visitor kv{&hp, seed_h};
  lane_visit(g, qx, qy, qz, kv);
  if (hp.size < hp.k && seed_h != cap_h) {
    hp.size = 0;
    kv.wh = cap_h;
    lane_visit(g, qx, qy, qz, kv);
  }
  const int m = hp.size;
  if (m == 0) {
    r2_out = u2f(cap_h);
    return 0;
  }
  const bool full = m == hp.k;
  SumVisitor<Add> sv{full ? hp.h(0) : cap_h, !full, 0u, 0u, LKEY_INF, &add};
  lane_visit(g, qx, qy, qz, sv);
  if (full) {
    uint32_t need = (uint32_t)hp.k - sv.n_less;  // >= 1: points at exactly D still to take
    if (need >= 1u && sv.n_tie >= 1u) {
      add(sv.tie_min, sv.D);
      uint32_t last = sv.tie_min;
      for (uint32_t r = 1; r < need; ++r) {
        TieVisitor tv{sv.D, last, LKEY_INF};
        lane_visit(g, qx, qy, qz, tv);
        if (tv.best == LKEY_INF) break;
        add(tv.best, sv.D);
        last = tv.best;
      }
    }
  }
  r2_out = full ? u2f(sv.D) : u2f(cap_h);
  return m;
}

template <bool CONE>
struct LaneGather2 {
  GatherMap gs, gv;
  LaneHeap32 hp;
  float hg_g;
  uint32_t nsurf, nvol;
  V3f prev_x;
  float prev_r;

  struct SurfAdd {
    const GatherMap* g;
    V3f n;
    float r;
    NSum3<float> s;
    __device__ void operator()(uint32_t idx, uint32_t d2) {
      V3f pw, dir;
      load_payload_l(*g, idx, pw, dir);
      if (dot(dir, n) < 0.0f) {
        float w = CONE ? cone_weight(u2f(d2), r) : 1.0f;
        nsum3_add(s, pw * w);
      }
    }
  };
  struct VolAdd {
    const GatherMap* g;
    V3f wo;
    float hg;
    NSum3<float> s;
    __device__ void operator()(uint32_t idx, uint32_t) {
      V3f pw, dir;
      load_payload_l(*g, idx, pw, dir);
      nsum3_add(s, pw * phase_hg(dot(dir, wo), hg));
    }
  };

  __device__ V3f surface(V3f x, V3f n, V3f albedo, V3f = V3f(), V3f = V3f()) {
    ++nsurf;
    const uint32_t caph = f2u(gs.rcap2);
    // the cone weight needs r before the sum: find the radius first
    hp.k = gs.k;
    hp.size = 0;
    Knn32Visitor kv{&hp, caph};
    lane_visit(gs, x.x, x.y, x.z, kv);
    const int m = hp.size;
    if (m == 0) return v3(0.0f, 0.0f, 0.0f);
    const bool full = m == gs.k;
    const float r2 = full ? u2f(hp.h(0)) : gs.rcap2;
    SurfAdd add{&gs, n, sqrt_(r2), {}};
    nsum3_init(add.s);
    SumVisitor<SurfAdd> sv{full ? hp.h(0) : caph, !full, 0u, 0u, LKEY_INF, &add};
    lane_visit(gs, x.x, x.y, x.z, sv);
    if (full) {
      uint32_t need = (uint32_t)gs.k - sv.n_less;
      if (need >= 1u && sv.n_tie >= 1u) {
        add(sv.tie_min, sv.D);
        uint32_t last = sv.tie_min;
        for (uint32_t r = 1; r < need; ++r) {
          TieVisitor tv{sv.D, last, LKEY_INF};
          lane_visit(gs, x.x, x.y, x.z, tv);
          if (tv.best == LKEY_INF) break;
          add(tv.best, sv.D);
          last = tv.best;
        }
      }
    }
    V3f sum = nsum3_get(add.s);
    V3f sc = v3(gs.scale[0], gs.scale[1], gs.scale[2]);
    return mul(mul(sum, albedo), sc) * surface_norm(r2, CONE);
  }

  __device__ V3f volume(V3f x, V3f wo, V3f = V3f()) {
    ++nvol;
    uint32_t caph = f2u(gv.rcap2), bh = caph;
    if (prev_r >= 0.0f) {
      float dx = x.x - prev_x.x, dy = x.y - prev_x.y, dz = x.z - prev_x.z;
      float b = (prev_r + sqrt_((dx * dx + dy * dy) + dz * dz)) * 1.0001f;
      float b2 = b * b;
      if (b2 < gv.rcap2) bh = f2u(b2);
    }
    VolAdd add{&gv, wo, hg_g, {}};
    nsum3_init(add.s);
    float r2;
    int m = knn32_gather(gv, hp, x.x, x.y, x.z, bh, caph, add, r2);
    prev_x = x;
    prev_r = m == gv.k ? sqrt_(r2) : -1.0f;
    if (m == 0) return v3(0.0f, 0.0f, 0.0f);
    V3f sum = nsum3_get(add.s);
    V3f sc = v3(gv.scale[0], gv.scale[1], gv.scale[2]);
    return mul(sum, sc) * volume_norm(r2);
  }
};

}  // namespace pm
