// pm_lane_knn3.h -- per-lane exact kNN radiance with a 2-byte heap.
//
// Gather mode 2 (pm_lane_knn2.h) keeps FP32 d^2 keys: 16 KB of LDS per wave
// at k = 64, 10 waves per compute unit.  Here the heap keeps a 16-bit,
// order-preserving bucket of d^2 (its FP32 bits >> 15: exponent and the top 7
// mantissa bits, i.e. buckets 2^-7 wide in relative d^2), 8 KB per wave, so
// the compute unit holds as many waves as its registers allow.
//   pass 1: the k smallest buckets under the search bound; the largest, Q, is
//           the bucket of the k-th neighbour;
//   pass 2: points in buckets below Q contribute; the points in bucket Q are
//           ordered by (d^2, index), and the (k - #below) smallest of them
//           contribute.  The k-th of those
//           gives the exact r_k.
// The neighbour set is the exact (d^2, index)-ordered kNN set, as in modes 1
// and 2.  A bucket holds the k-th neighbour and on average 1.5 k 2^-7 < 1
// other point at k = 64, so the two kept keys nearly always suffice; when
// more bucket points are needed, further exact passes pick them in order.
#pragma once

#include "pm_lane_knn2.h"

namespace pm {

__device__ __forceinline__ uint32_t dq16(uint32_t d2bits) { return d2bits >> 15; }
__device__ __forceinline__ uint32_t dq16_top(uint32_t q) { return ((q + 1u) << 15) - 1u; }  // largest d^2 bits in q

struct LaneHeap16 {
  uint16_t* H;
  int lane;
  int k;
  int size;
  __device__ uint32_t h(int j) const { return H[j * 64 + lane]; }
  __device__ void set(int j, uint32_t v) { H[j * 64 + lane] = (uint16_t)v; }
  __device__ void push(uint32_t v) {
    int j = size++;
    while (j > 0) {
      int p = (j - 1) >> 1;
      uint32_t pv = h(p);
      if (pv >= v) break;
      set(j, pv);
      j = p;
    }
    set(j, v);
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
      set(j, cv);
      j = c;
    }
    set(j, v);
  }
};

// Pass 1: the k smallest buckets among points with d^2 <= bound.
struct Knn16Visitor {
  LaneHeap16* hp;
  uint32_t bound_h;  // exact search bound (FP32 bits), used until the heap is full
  uint32_t root;     // bucket at the heap root once full
  bool full;
  __device__ uint32_t bound() const { return full ? dq16_top(root) : bound_h; }
  template <int NB>
  __device__ void points(const float4*, const uint32_t* dh, int cnt) {
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < cnt) {
        const uint32_t d = dh[j];
        if (!full) {
          if (d <= bound_h) {
            hp->push(dq16(d));
            if (hp->size == hp->k) {
              full = true;
              root = hp->h(0);
            }
          }
        } else if (dq16(d) < root) {
          hp->replace_root(dq16(d));
          root = hp->h(0);
        }
      }
    }
  }
};

// Pass 2: points below bucket Q (or, inclusive, with d^2 <= cap) go to `add`;
// of the bucket-Q points the two smallest (d^2, index) keys are kept.
template <class Add>
struct Sum16Visitor {
  uint32_t Q;        // bucket of the k-th neighbour
  uint32_t cap_h;    // inclusive mode: every point with d^2 <= cap_h
  bool inclusive;
  uint32_t n_less;
  uint32_t d0, i0, d1, i1;  // (d0, i0) < (d1, i1), LKEY_INF when empty
  Add* add;
  __device__ uint32_t bound() const { return inclusive ? cap_h : dq16_top(Q); }
  __device__ void keep(uint32_t d, uint32_t idx) {
    if (key_less(d, idx, d0, i0)) {
      d1 = d0;
      i1 = i0;
      d0 = d;
      i0 = idx;
    } else if (key_less(d, idx, d1, i1)) {
      d1 = d;
      i1 = idx;
    }
  }
  template <int NB>
  __device__ void points(const float4* p, const uint32_t* dh, int cnt) {
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < cnt) {
        const uint32_t d = dh[j];
        const uint32_t idx = f2u(p[j].w);
        const bool take = inclusive ? d <= cap_h : dq16(d) < Q;
        if (take) {
          (*add)(idx, d);
          ++n_less;
        } else if (!inclusive && dq16(d) == Q) {
          keep(d, idx);
        }
      }
    }
  }
};

// Smallest (d^2, index) key in bucket Q above (last_d, last_i) (overflow path).
struct Bucket16Visitor {
  uint32_t Q, last_d, last_i, best_d, best_i;
  __device__ uint32_t bound() const { return dq16_top(Q); }
  template <int NB>
  __device__ void points(const float4* p, const uint32_t* dh, int cnt) {
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < cnt && dq16(dh[j]) == Q) {
        const uint32_t idx = f2u(p[j].w);
        if (key_less(last_d, last_i, dh[j], idx) && key_less(dh[j], idx, best_d, best_i)) {
          best_d = dh[j];
          best_i = idx;
        }
      }
    }
  }
};

// kNN radiance gather: pass 1 with the bounded-then-capped retry of the
// volume gather, pass 2, bucket selection.  add(idx, d2 bits) is called once
// per neighbour; returns m and sets r2_out (k-th d^2, or the cap).
template <int NB, class Add>
__device__ __forceinline__ int knn16_gather(const GatherMap& g, LaneHeap16& hp, float qx, float qy, float qz,
                                            uint32_t seed_h, uint32_t cap_h, Add& add, float& r2_out) {
  hp.k = g.k;
  hp.size = 0;
  Knn16Visitor kv{&hp, seed_h, 0u, false};
  lane_visit<NB>(g, qx, qy, qz, kv);
  if (!kv.full && seed_h != cap_h) {
    hp.size = 0;
    kv.bound_h = cap_h;
    lane_visit<NB>(g, qx, qy, qz, kv);
  }
  const int m = hp.size;
  r2_out = u2f(cap_h);
  if (m == 0) return 0;
  Sum16Visitor<Add> sv{kv.root, cap_h, !kv.full, 0u, LKEY_INF, LKEY_INF, LKEY_INF, LKEY_INF, &add};
  lane_visit<NB>(g, qx, qy, qz, sv);
  if (!kv.full) return m;
  const uint32_t need = (uint32_t)g.k - sv.n_less;  // >= 1: bucket points still to take
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
  for (uint32_t r = 2; r < need; ++r) {  // more than two bucket points needed: exact passes
    Bucket16Visitor bv{sv.Q, kd, ki, LKEY_INF, LKEY_INF};
    lane_visit<NB>(g, qx, qy, qz, bv);
    if (bv.best_i == LKEY_INF) break;
    add(bv.best_i, bv.best_d);
    kd = bv.best_d;
    ki = bv.best_i;
  }
  if (kd != LKEY_INF) r2_out = u2f(kd);
  return m;
}

template <bool CONE, int NB = 8>
struct LaneGather3 {
  GatherMap gs, gv;
  LaneHeap16 hp;
  float hg_g;
  uint32_t nsurf, nvol;
  V3f prev_x;
  float prev_r;

  // Surface sums: filtered power, and power x distance for the cone weight
  // 1 - d / (1.1 r), applied once r is known.
  struct SurfAdd {
    const GatherMap* g;
    V3f n;
    NSum3<float> s0, s1;
    __device__ void operator()(uint32_t idx, uint32_t d2) {
      V3f pw, dir;
      load_payload_l(*g, idx, pw, dir);
      if (dot(dir, n) < 0.0f) {
        nsum3_add(s0, pw);
        if (CONE) nsum3_add(s1, pw * sqrt_(u2f(d2)));
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
    int m = knn16_gather<NB>(gv, hp, x.x, x.y, x.z, bh, caph, add, r2);
    prev_x = x;
    prev_r = m == gv.k ? sqrt_(r2) : -1.0f;
    if (m == 0) return v3(0.0f, 0.0f, 0.0f);
    V3f sum = nsum3_get(add.s);
    V3f sc = v3(gv.scale[0], gv.scale[1], gv.scale[2]);
    return mul(sum, sc) * volume_norm(r2);
  }
};

}  // namespace pm
