// pm_lane_knn.h -- per-lane exact kNN (LDS max-heap, stackless traversal)
// and the per-lane gather object used by the lane-mode render and the cache
// kernels.  Device-only; see pm_gather_lane.hip for the design notes.
#pragma once

#include "pm_kernels.h"
#include "pm_wave.h"

namespace pm {


static constexpr uint32_t LKEY_INF = 0xffffffffu;

__device__ __forceinline__ bool kgt(uint32_t ah, uint32_t al, uint32_t bh, uint32_t bl) {
  return ah > bh || (ah == bh && al > bl);
}

// Per-lane max-heap in LDS: slot j of this lane at H[j * 64 + lane].
struct LaneHeap {
  uint32_t* H;
  uint32_t* Lo;
  int lane;
  int k;
  int size;
  __device__ uint32_t& h(int j) { return H[j * 64 + lane]; }
  __device__ uint32_t& l(int j) { return Lo[j * 64 + lane]; }
  __device__ void push(uint32_t nh, uint32_t nl) {
    int j = size++;
    while (j > 0) {
      int p = (j - 1) >> 1;
      uint32_t ph = h(p), pl = l(p);
      if (!kgt(nh, nl, ph, pl)) break;
      h(j) = ph;
      l(j) = pl;
      j = p;
    }
    h(j) = nh;
    l(j) = nl;
  }
  __device__ void replace_root(uint32_t nh, uint32_t nl) {
    int j = 0;
    for (;;) {
      int c = 2 * j + 1;
      if (c >= k) break;
      uint32_t ch = h(c), cl = l(c);
      if (c + 1 < k) {
        uint32_t dh = h(c + 1), dl = l(c + 1);
        if (kgt(dh, dl, ch, cl)) {
          ++c;
          ch = dh;
          cl = dl;
        }
      }
      if (!kgt(ch, cl, nh, nl)) break;
      h(j) = ch;
      l(j) = cl;
      j = c;
    }
    h(j) = nh;
    l(j) = nl;
  }
};

// Squared distance from q to a node's tight box, as a lower bound of the
// squared distance to any of its points: with the operation order of dist2,
// and FP32 rounding monotonic, no point in the box has a smaller computed d^2.
__device__ __forceinline__ float box_d2(V3f q, float4 mn, float4 mx) {
  const float bx = fmaxf(fmaxf(mn.x - q.x, q.x - mx.x), 0.0f);
  const float by = fmaxf(fmaxf(mn.y - q.y, q.y - mx.y), 0.0f);
  const float bz = fmaxf(fmaxf(mn.z - q.z, q.z - mx.z), 0.0f);
  return (bx * bx + by * by) + bz * bz;
}

// Exact kNN of q for this lane; the heap holds the result (unordered), its
// size is returned.  bound_h: see wave_knn in pm_gather.hip.  When the map has
// node boxes, a far child or a leaf whose box lies beyond the current bound is
// skipped: none of its points can enter the heap (box_d2 is a lower bound of
// their d^2, and the heap takes (d^2, index) keys below its root).
__device__ __forceinline__ int lane_knn(const GatherMap& g, float qx, float qy, float qz, uint32_t bound_h,
                                        LaneHeap& hp) {
  hp.size = 0;
  hp.k = g.k;
  const uint32_t n = g.n, nl = g.nl;
  if (n == 0) return 0;
  uint32_t wh = bound_h, wl = LKEY_INF;
  uint32_t h = 0, lo = 0, hi = n, from = 0;
  bool down = true, done = false;
  // "while-while": every lane first walks to its next leaf (cheap node steps),
  // then the lanes at a leaf test its points together, then insert the
  // passing candidates together, so the costly parts run with most lanes on.
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
      // going up from node `from`
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
      const uint32_t farc = nearc == 2u * p + 1u ? 2u * p + 2u : 2u * p + 1u;
      if (from == nearc && diff * diff <= u2f(wh) &&
          (!g.bmin || box_d2(v3(qx, qy, qz), g.bmin[farc], g.bmax[farc]) <= u2f(wh))) {
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
    // leaf [lo, hi), at most 64 points, in batches of 8: all 8 loads are
    // issued before any is used, then the passing keys are inserted.
    const bool reach = !g.bmin || box_d2(v3(qx, qy, qz), g.bmin[h], g.bmax[h]) <= u2f(wh);
    for (uint32_t b0 = lo; reach && b0 < hi; b0 += 8u) {
      float4 p[8];
      const float4* q = g.tpts + b0;  // padded: entries past hi are read, not used
#pragma unroll
      for (int j = 0; j < 8; ++j) p[j] = q[j];
      uint32_t dh[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) dh[j] = f2u(dist2(qx, qy, qz, p[j].x, p[j].y, p[j].z));
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        uint32_t di = f2u(p[j].w);
        if (b0 + (uint32_t)j < hi && key_less(dh[j], di, wh, wl)) {
          if (hp.size < hp.k) {
            hp.push(dh[j], di);
            if (hp.size == hp.k) {
              wh = hp.h(0);
              wl = hp.l(0);
            }
          } else {
            hp.replace_root(dh[j], di);
            wh = hp.h(0);
            wl = hp.l(0);
          }
        }
      }
    }
    from = h;
    down = false;
  }
  return hp.size;
}

__device__ __forceinline__ void load_payload_l(const GatherMap& g, uint32_t idx, V3f& pw, V3f& dir) {
  if (g.bf16) {
    uint4 w = ((const uint4*)g.payload)[idx];
    pw = v3(bf16_to_f32((uint16_t)(w.x & 0xffffu)), bf16_to_f32((uint16_t)(w.x >> 16)),
            bf16_to_f32((uint16_t)(w.y & 0xffffu)));
    dir = v3(bf16_to_f32((uint16_t)(w.y >> 16)), bf16_to_f32((uint16_t)(w.z & 0xffffu)),
             bf16_to_f32((uint16_t)(w.z >> 16)));
  } else {
    const float4* q = (const float4*)g.payload;
    float4 a = q[2 * (size_t)idx];
    float2 b = *(const float2*)&q[2 * (size_t)idx + 1];
    pw = v3(a.x, a.y, a.z);
    dir = v3(a.w, b.x, b.y);
  }
}

// The same from the tree-ordered copy (tree position pos).
__device__ __forceinline__ void load_payload_tree(const GatherMap& g, uint32_t pos, V3f& pw, V3f& dir) {
  if (g.bf16) {
    uint4 w = ((const uint4*)g.tpayload)[pos];
    pw = v3(bf16_to_f32((uint16_t)(w.x & 0xffffu)), bf16_to_f32((uint16_t)(w.x >> 16)),
            bf16_to_f32((uint16_t)(w.y & 0xffffu)));
    dir = v3(bf16_to_f32((uint16_t)(w.y >> 16)), bf16_to_f32((uint16_t)(w.z & 0xffffu)),
             bf16_to_f32((uint16_t)(w.z >> 16)));
  } else {
    const float4* q = (const float4*)g.tpayload;
    float4 a = q[2 * (size_t)pos];
    float2 b = *(const float2*)&q[2 * (size_t)pos + 1];
    pw = v3(a.x, a.y, a.z);
    dir = v3(a.w, b.x, b.y);
  }
}

template <bool CONE>
struct LaneGather {
  GatherMap gs, gv;
  LaneHeap hp;
  float hg_g;
  uint32_t nsurf, nvol;
  V3f prev_x;
  float prev_r;

  __device__ V3f surface(V3f x, V3f n, V3f albedo, V3f = V3f(), V3f = V3f()) {
    ++nsurf;
    int m = lane_knn(gs, x.x, x.y, x.z, f2u(gs.rcap2), hp);
    if (m == 0) return v3(0.0f, 0.0f, 0.0f);
    float r2 = m == gs.k ? u2f(hp.h(0)) : gs.rcap2;
    float r = sqrt_(r2);
    NSum3<float> s;
    nsum3_init(s);
    for (int j = 0; j < m; ++j) {
      V3f pw, dir;
      load_payload_l(gs, hp.l(j), pw, dir);
      if (dot(dir, n) < 0.0f) {
        float w = CONE ? cone_weight(u2f(hp.h(j)), r) : 1.0f;
        nsum3_add(s, pw * w);
      }
    }
    V3f sum = nsum3_get(s);
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
    int m = lane_knn(gv, x.x, x.y, x.z, bh, hp);
    if (m < gv.k && bh != caph) m = lane_knn(gv, x.x, x.y, x.z, caph, hp);
    prev_x = x;
    prev_r = m == gv.k ? sqrt_(u2f(hp.h(0))) : -1.0f;
    if (m == 0) return v3(0.0f, 0.0f, 0.0f);
    float r2 = m == gv.k ? u2f(hp.h(0)) : gv.rcap2;
    NSum3<float> s;
    nsum3_init(s);
    for (int j = 0; j < m; ++j) {
      V3f pw, dir;
      load_payload_l(gv, hp.l(j), pw, dir);
      nsum3_add(s, pw * phase_hg(dot(dir, wo), hg_g));
    }
    V3f sum = nsum3_get(s);
    V3f sc = v3(gv.scale[0], gv.scale[1], gv.scale[2]);
    return mul(sum, sc) * volume_norm(r2);
  }
};

}  // namespace pm
