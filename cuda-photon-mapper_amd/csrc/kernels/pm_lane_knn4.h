// pm_lane_knn4.h -- the 2-byte-heap gather (pm_lane_knn3.h) with the march's
// volume queries answered one step late, so that the exact second pass of
// step j and the first pass of step j + 1 share one traversal.
//
// Consecutive march points are one step apart and their neighbour balls
// overlap almost entirely, so the two passes visit nearly the same leaves.
// The shared traversal descends toward the new point and visits a far child
// when it can hold points of either ball; each leaf point is tested against
// both queries.  The neighbour sets, r_k and the estimates are those of the
// two-pass gather; only the order of the traversal work changes.  The first
// pass of a point is seeded with the bound (r_prev + |x - x_prev|) where r_prev
// is the top of the previous point's k-th-neighbour bucket (its exact r_k is
// only known one step later), which is still an upper bound.
#pragma once

#include "pm_lane_knn3.h"

namespace pm {

// Stackless traversal toward q that asks the visitor whether a far child can
// hold wanted points: V::far(axis, split) and V::points<NB>(p, cnt).
template <int NB, class V>
__device__ __forceinline__ void lane_visit_dual(const GatherMap& g, float qx, float qy, float qz, V& v) {
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
      uint32_t nearc = qa - nd.split < 0.0f ? 2u * p + 1u : 2u * p + 2u;
      if (from == nearc && v.far(a, nd.split)) {
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
      const uint32_t left = hi - b0;
      v.template points<NB>(p, left < (uint32_t)NB ? (int)left : NB);
    }
    from = h;
    down = false;
  }
}

__device__ __forceinline__ float axis_of(V3f q, int a) { return a == 0 ? q.x : (a == 1 ? q.y : q.z); }

// First pass of the new point (heap) and second pass of the pending point
// (sums) over one traversal.
template <class Add>
struct DualVisitor {
  V3f qn, qp;             // new point, pending point
  Knn16Visitor* kv;
  Sum16Visitor<Add>* sv;
  // The traversal is near-first for the new point; the far child is visited
  // when the new point's ball crosses the plane, or the pending point lies on
  // the far side, or its ball crosses the plane.
  __device__ bool far(int a, float split) const {
    const float dn = axis_of(qn, a) - split, dp = axis_of(qp, a) - split;
    const bool same_side = (dn < 0.0f) == (dp < 0.0f);
    return dn * dn <= u2f(kv->bound()) || !same_side || dp * dp <= u2f(sv->bound());
  }
  template <int NB>
  __device__ void points(const float4* p, int cnt) {
    uint32_t dn[NB], dp[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      dn[j] = f2u(dist2(qn.x, qn.y, qn.z, p[j].x, p[j].y, p[j].z));
      dp[j] = f2u(dist2(qp.x, qp.y, qp.z, p[j].x, p[j].y, p[j].z));
    }
    kv->template points<NB>(p, dn, cnt);
    sv->template points<NB>(p, dp, cnt);
  }
};

template <bool CONE, int NB = 4>
struct LaneGather4 {
  static constexpr bool kDeferredVolume = true;
  GatherMap gs, gv;
  LaneHeap16 hp;
  float hg_g;
  uint32_t nsurf, nvol;
  V3f prev_x;
  float prev_r;  // upper bound of the previous point's r_k (-1: none)
  // the queued volume point
  bool pend;
  bool pend_full;
  int pend_m;
  uint32_t pend_Q;
  V3f pend_x, pend_wo, pend_w;

  using SurfAdd = typename LaneGather3<CONE, NB>::SurfAdd;
  using VolAdd = typename LaneGather3<CONE, NB>::VolAdd;

  __device__ V3f surface(V3f x, V3f n, V3f albedo, V3f = V3f(), V3f = V3f()) {
    ++nsurf;
    const uint32_t caph = f2u(gs.rcap2);
    SurfAdd add;
    add.g = &gs;
    add.n = n;
    nsum3_init(add.s0);
    nsum3_init(add.s1);
    float r2;
    int m = knn16_gather<NB>(gs, hp, x.x, x.y, x.z, caph, caph, add, r2);
    if (m == 0) return v3(0.0f, 0.0f, 0.0f);
    V3f sum = nsum3_get(add.s0);
    if (CONE) sum = sum - nsum3_get(add.s1) * (1.0f / (1.1f * sqrt_(r2)));
    V3f sc = v3(gs.scale[0], gs.scale[1], gs.scale[2]);
    return mul(mul(sum, albedo), sc) * surface_norm(r2, CONE);
  }

  // Bucket selection of the pending point after its second pass: adds the
  // (k - #below) smallest bucket keys (extra exact passes beyond two) and
  // returns w * value.
  template <class SV, class A>
  __device__ V3f finish_pending(SV& sv, A& add) {
    const uint32_t cap_h = f2u(gv.rcap2);
    float r2 = u2f(cap_h);
    if (pend_full) {
      const uint32_t need = (uint32_t)gv.k - sv.n_less;
      uint32_t kd = LKEY_INF, ki = LKEY_INF;
      if (sv.d0 != LKEY_INF) {
        add(sv.i0, sv.d0);
        kd = sv.d0;
        ki = sv.i0;
      }
      if (need >= 2u && sv.d1 != LKEY_INF) {
        add(sv.i1, sv.d1);
        kd = sv.d1;
        ki = sv.i1;
      }
      for (uint32_t r = 2; r < need; ++r) {
        Bucket16Visitor bv{pend_Q, kd, ki, LKEY_INF, LKEY_INF};
        lane_visit<NB>(gv, pend_x.x, pend_x.y, pend_x.z, bv);
        if (bv.best_i == LKEY_INF) break;
        add(bv.best_i, bv.best_d);
        kd = bv.best_d;
        ki = bv.best_i;
      }
      if (kd != LKEY_INF) r2 = u2f(kd);
    }
    V3f sum = nsum3_get(add.s);
    V3f sc = v3(gv.scale[0], gv.scale[1], gv.scale[2]);
    return mul(pend_w, mul(sum, sc) * volume_norm(r2));
  }

  // First pass of x alone (and the capped retry); the point becomes pending.
  __device__ void first_pass_alone(V3f x, uint32_t seed_h, uint32_t cap_h, Knn16Visitor& kv) {
    lane_visit<NB>(gv, x.x, x.y, x.z, kv);
    if (!kv.full && seed_h != cap_h) {
      hp.size = 0;
      kv.bound_h = cap_h;
      lane_visit<NB>(gv, x.x, x.y, x.z, kv);
    }
  }

  __device__ void queue(V3f x, V3f wo, V3f w, const Knn16Visitor& kv) {
    pend = true;
    pend_full = kv.full;
    pend_m = hp.size;
    pend_Q = kv.root;
    pend_x = x;
    pend_wo = wo;
    pend_w = w;
    prev_x = x;
    prev_r = kv.full ? sqrt_(u2f(dq16_top(kv.root))) : -1.0f;
  }

  __device__ bool volume_defer(V3f x, V3f wo, V3f w, V3f& c) {
    ++nvol;
    const uint32_t caph = f2u(gv.rcap2);
    uint32_t bh = caph;
    if (prev_r >= 0.0f) {
      float dx = x.x - prev_x.x, dy = x.y - prev_x.y, dz = x.z - prev_x.z;
      float b = (prev_r + sqrt_((dx * dx + dy * dy) + dz * dz)) * 1.0001f;
      float b2 = b * b;
      if (b2 < gv.rcap2) bh = f2u(b2);
    }
    hp.k = gv.k;
    hp.size = 0;
    Knn16Visitor kv{&hp, bh, 0u, false};
    if (!pend) {
      first_pass_alone(x, bh, caph, kv);
      queue(x, wo, w, kv);
      return false;
    }
    VolAdd add{&gv, pend_wo, hg_g, {}};
    nsum3_init(add.s);
    Sum16Visitor<VolAdd> sv{pend_Q, caph, !pend_full, 0u, LKEY_INF, LKEY_INF, LKEY_INF, LKEY_INF, &add};
    if (pend_m > 0) {
      DualVisitor<VolAdd> dv{x, pend_x, &kv, &sv};
      lane_visit_dual<NB>(gv, x.x, x.y, x.z, dv);
      if (!kv.full && bh != caph) {
        hp.size = 0;
        kv.bound_h = caph;
        lane_visit<NB>(gv, x.x, x.y, x.z, kv);
      }
      c = finish_pending(sv, add);
    } else {
      first_pass_alone(x, bh, caph, kv);
      c = mul(pend_w, v3(0.0f, 0.0f, 0.0f));
    }
    queue(x, wo, w, kv);
    return true;
  }

  __device__ bool volume_flush(V3f, V3f& c) {
    if (!pend) return false;
    pend = false;
    if (pend_m == 0) {
      c = mul(pend_w, v3(0.0f, 0.0f, 0.0f));
      return true;
    }
    const uint32_t caph = f2u(gv.rcap2);
    VolAdd add{&gv, pend_wo, hg_g, {}};
    nsum3_init(add.s);
    Sum16Visitor<VolAdd> sv{pend_Q, caph, !pend_full, 0u, LKEY_INF, LKEY_INF, LKEY_INF, LKEY_INF, &add};
    lane_visit<NB>(gv, pend_x.x, pend_x.y, pend_x.z, sv);
    c = finish_pending(sv, add);
    return true;
  }
};

}  // namespace pm
