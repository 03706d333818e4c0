// pm_lane_knn5.h -- the deferred-volume gather (pm_lane_knn4.h) with the
// first pass of a march point done by counting instead of a heap.
//
// The first pass only has to find Q, the 16-bit bucket (FP32 bits of d^2
// >> 15) of the k-th neighbour; the second pass then takes the exact
// neighbour set from Q.  Along a march the k-th-neighbour radius changes little
// from one step to the next, so the first pass of a point counts its
// neighbours in a window of 256 consecutive buckets, [Q' - 160, Q' + 96) around
// the previous point's bucket Q' (a factor 2^(5/8) down and 2^(3/8) up in
// radius):
//   * points in buckets below the window increment a register counter;
//   * points in the window increment their bucket's 4-bit counter in LDS (256
//     counters per lane, 8 KB per wave, in the LDS of the heap);
//   * points above the window are outside the search bound, which is the top
//     of the window (or the radius cap), so the traversal never visits them.
// Q is the first bucket at which the running count reaches k.  When the window
// reaches the cap and holds fewer than k points, the point has fewer than k
// neighbours within the cap, and its count is known.  A point whose k-th
// neighbour falls outside the window, or one of whose counters wrapped (more
// than 15 points in a bucket), runs the heap pass of pm_lane_knn4.h instead.
// Either way Q is the bucket that the heap pass finds, so neighbour sets, r_k
// and estimates are those of the other gather modes.
// Counting costs one LDS add per point in the window and no data-dependent
// loop, where the heap's insertions run sift loops that the whole wave waits
// through for the few lanes that insert.
// Queries gathered on their own (the march's surface hits with the box
// estimate, and the radiance cache's exact values) have no previous point to
// seed a window: they run the heap pass and the second pass on the
// speculative traversal with node boxes (surface_lone, SPEC_HEAP).
#pragma once

#include "pm_lane_knn4.h"

namespace pm {

typedef __attribute__((address_space(3))) uint32_t lds_u32;

// First pass by counting in the bucket window [K0, K0 + 256).
struct HistVisitor {
  static constexpr uint32_t kBuckets = 256;  // 4-bit counters: 32 dwords per lane, 8 KB per wave
  lds_u32* H;        // this wave's counters: dword w of lane l at H[w * 64 + l], 8 buckets per dword
  int lane;
  uint32_t K0;       // first bucket of the window
  uint32_t top_h;    // search bound (FP32 bits of d^2): at most the top of the window
  uint32_t below;    // points under the bound in buckets below K0
  uint32_t inwin;    // points in the window
  __device__ uint32_t bound() const { return top_h; }
  __device__ void clear() {
#pragma unroll
    for (uint32_t w = 0; w < kBuckets / 8; ++w) H[w * 64 + lane] = 0u;
  }
  template <int NB>
  __device__ void points(const uint32_t* dh, int cnt) {
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const uint32_t d = dh[j];
      if (j < cnt && d <= top_h) {
        const uint32_t q = dq16(d);
        if (q < K0) {
          ++below;
        } else {
          const uint32_t b = q - K0;  // < kBuckets: d <= top_h <= top of bucket K0 + kBuckets - 1
          __hip_atomic_fetch_add(&H[(b >> 3) * 64u + (uint32_t)lane], 1u << (4u * (b & 7u)), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WAVEFRONT);
          ++inwin;
        }
      }
    }
  }
  // The bucket Q of the k-th neighbour under the bound; false when fewer than
  // k points lie under it, when k points lie below the window, or when a
  // counter wrapped.  A counter that passes 15 carries into the next one (or
  // out of the dword), which lowers the sum of all counters by 15 (or 16) and
  // can never raise it, so the counters are exact exactly when their sum is
  // the register count `inwin`.
  __device__ bool select(uint32_t k, uint32_t& Q) const {
    if (below >= k || below + inwin < k) return false;
    const uint32_t need = k - below;
    uint32_t cum = 0u, wq = 0u, vq = 0u, before = 0u;
    bool found = false;
#pragma unroll
    for (uint32_t w = 0; w < kBuckets / 8; ++w) {
      const uint32_t v = H[w * 64 + lane];
      // the 8 counters summed pairwise into 4 bytes (each <= 30), then the
      // bytes (sum <= 120, so no byte carries)
      const uint32_t b = (v & 0x0F0F0F0Fu) + ((v >> 4) & 0x0F0F0F0Fu);
      const uint32_t c = cum + ((b * 0x01010101u) >> 24);
      if (!found && c >= need) {
        found = true;
        wq = w;
        vq = v;
        before = cum;
      }
      cum = c;
    }
    if (cum != inwin) return false;
    // first counter of dword wq at which the running count reaches `need`
    uint32_t run = before, s = 0u;
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      run += (vq >> (4 * i)) & 15u;
      s += run < need ? 1u : 0u;
    }
    Q = K0 + 8u * wq + s;
    return found;
  }
};

// Volume contributions, by record index (the tie picks of the exact
// selection) or by tree position (the second pass, which reads the payload
// beside the leaf points it has just read).
struct VolAddT {
  const GatherMap* g;
  V3f wo;
  float hg;
  NSum3<float> s;
  __device__ void add(V3f pw, V3f dir) { nsum3_add(s, pw * phase_hg(dot(dir, wo), hg)); }
  __device__ void operator()(uint32_t idx, uint32_t) {
    V3f pw, dir;
    load_payload_l(*g, idx, pw, dir);
    add(pw, dir);
  }
  __device__ void tree(uint32_t pos) {
    V3f pw, dir;
    load_payload_tree(*g, pos, pw, dir);
    add(pw, dir);
  }
};

// Surface contributions without the cone filter (power of the photons that
// arrived from the query's side), by record index or by tree position.
struct SurfAddT {
  const GatherMap* g;
  V3f n;
  NSum3<float> s;
  __device__ void add(V3f pw, V3f dir) {
    if (dot(dir, n) < 0.0f) nsum3_add(s, pw);
  }
  __device__ void operator()(uint32_t idx, uint32_t) {
    V3f pw, dir;
    load_payload_l(*g, idx, pw, dir);
    add(pw, dir);
  }
  __device__ void tree(uint32_t pos) {
    V3f pw, dir;
    load_payload_tree(*g, pos, pw, dir);
    add(pw, dir);
  }
};

// Second pass with the contributions of a batch compacted per lane: a lane's
// neighbours among the batch's points are gathered first, then the
// contribution code runs max-over-lanes times instead of once per point
// position that any lane takes (same points, same order).  The batch's points
// are tree positions base, base + 1, ...; contributions come from the
// tree-ordered payload.
template <class Add>
struct Sum16CVisitor : Sum16Visitor<Add> {
  template <int NB>
  __device__ void points(const float4* p, const uint32_t* dh, int cnt, uint32_t base) {
    // the tail below names sel[2] and sel[3] directly (no dynamic register
    // indexing), which covers batches of at most 4 points
    static_assert(NB <= 4, "Sum16CVisitor handles leaf batches of at most 4 points");
    uint32_t sel[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) sel[j] = 0u;
    uint32_t nt = 0u;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const uint32_t d = dh[j];
      const bool in = j < cnt;
      const bool take = in && (this->inclusive ? d <= this->cap_h : dq16(d) < this->Q);
#pragma unroll
      for (int t = 0; t <= j; ++t) sel[t] = take && nt == (uint32_t)t ? base + (uint32_t)j : sel[t];
      nt += take ? 1u : 0u;
      if (in && !take && !this->inclusive && dq16(d) == this->Q) this->keep(d, f2u(p[j].w));
    }
    this->n_less += nt;
    if (nt == 0u) return;
    // the first two payloads are loaded together, then summed in order
    V3f pw0, dir0, pw1, dir1;
    load_payload_tree(*this->add->g, sel[0], pw0, dir0);
    if (nt > 1u) load_payload_tree(*this->add->g, sel[1], pw1, dir1);
    this->add->add(pw0, dir0);
    if (nt > 1u) this->add->add(pw1, dir1);
    for (uint32_t t = 2; t < nt; ++t) this->add->tree(t == 2u ? sel[2] : sel[3]);
  }
};

// lane_visit_dual of pm_lane_knn4.h for visitors whose far-child test does not
// change during the traversal (fixed search bounds), with speculative
// traversal (T. Aila and S. Laine, "Understanding the Efficiency of Ray
// Traversal on GPUs", HPG 2009): a lane that reaches a leaf while other lanes
// are still descending keeps that leaf for later and walks on toward its next
// leaf, instead of waiting; the leaves are processed once every active lane
// holds one.  Each lane visits the same leaves in the same order as without
// speculation, because the far-child test does not depend on what the leaves
// held.  When the map has node boxes, a far child and a leaf are entered only
// if the visitor's box test (V::near_box) passes: the points skipped are all
// outside every search bound.
// Node reads of the traversal: from the map's top-of-tree tile in LDS when
// TOP and the node is in it, else from HBM.  The values are the same.
template <bool TOP>
__device__ __forceinline__ KdNode tree_node(const GatherMap& g, uint32_t h) {
  if constexpr (TOP) {
    if (h < g.top_count) return g.top_nodes[h];
  }
  return g.nodes[h];
}
template <bool TOP>
__device__ __forceinline__ uint2 tree_range(const GatherMap& g, uint32_t h) {
  if constexpr (TOP) {
    if (h < g.top_count) return g.top_nrange[h];
  }
  return g.nrange[h];
}
// The visitor's box test on node h (true when the map has no boxes).
template <bool TOP, class V>
__device__ __forceinline__ bool tree_box_near(const GatherMap& g, uint32_t h, const V& v) {
  if (!g.bmin) return true;
  if constexpr (TOP) {
    if (h < g.top_count) return v.near_box(g.top_bmin[h], g.top_bmax[h]);
  }
  return v.near_box(g.bmin[h], g.bmax[h]);
}

// Stages the first `levels` levels of g's internal nodes (at most nl - 1
// nodes) into `lds` (top_tile_bytes(levels) bytes, 16-byte aligned) with the
// 64 lanes of the wave, and points g's top tile at them.  Every lane of the
// wave must call it; the copies are visible to the wave when it returns.
__device__ __forceinline__ void stage_top_tile(GatherMap& g, unsigned char* lds, int levels, int lane) {
  const uint32_t cap = ((uint32_t)1 << levels) - 1u;
  const uint32_t ni = g.nl > 0u ? g.nl - 1u : 0u;
  const uint32_t c = ni < cap ? ni : cap;
  float4* bmn = (float4*)lds;
  float4* bmx = bmn + cap;
  KdNode* nd = (KdNode*)(bmx + cap);
  uint2* nr = (uint2*)(nd + cap);
  // the wave's earlier reads of this region (a previous tile) come first
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (uint32_t i = (uint32_t)lane; i < c; i += 64u) {
    nd[i] = g.nodes[i];
    nr[i] = g.nrange[i];
    if (g.bmin) {
      bmn[i] = g.bmin[i];
      bmx[i] = g.bmax[i];
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  g.top_nodes = nd;
  g.top_nrange = nr;
  g.top_bmin = bmn;
  g.top_bmax = bmx;
  g.top_count = c;
}

template <int NB, bool TOP = false, class V>
__device__ __forceinline__ void lane_visit_dual_spec(const GatherMap& g, float qx, float qy, float qz, V& v) {
  const uint32_t n = g.n, nl = g.nl;
  if (n == 0) return;
  uint32_t h = 0, lo = 0, hi = n, from = 0;
  uint32_t plo = 0, phi = 0;  // the kept leaf's point range
  bool down = true, done = false, kept = false;
  for (;;) {
    for (;;) {
      if (!done && !kept && down && h + 1u >= nl) {
        // keep this leaf (unless its box is out of reach) and walk on
        if (tree_box_near<TOP>(g, h, v)) {
          plo = lo;
          phi = hi;
          kept = true;
        }
        from = h;
        down = false;
      }
      if (ballot(!done && !kept) == 0ull) break;
      const bool step = !done && !(down && h + 1u >= nl);
      if (!step) continue;
      if (down) {
        KdNode nd = tree_node<TOP>(g, h);
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
      KdNode nd = tree_node<TOP>(g, p);
      int a = kd_axis(nd.meta);
      float qa = a == 0 ? qx : (a == 1 ? qy : qz);
      uint32_t nearc = qa - nd.split < 0.0f ? 2u * p + 1u : 2u * p + 2u;
      const uint32_t farc = nearc == 2u * p + 1u ? 2u * p + 2u : 2u * p + 1u;
      if (from == nearc && v.far(a, nd.split) && tree_box_near<TOP>(g, farc, v)) {
        uint2 r = tree_range<TOP>(g, p);
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
    if (ballot(kept) == 0ull) break;
    if (!kept) continue;
    for (uint32_t b0 = plo; b0 < phi; b0 += (uint32_t)NB) {
      float4 p[NB];
      const float4* q = g.tpts + b0;  // padded: entries past phi are read, not used
#pragma unroll
      for (int j = 0; j < NB; ++j) p[j] = q[j];
      const uint32_t left = phi - b0;
      v.template points<NB>(p, left < (uint32_t)NB ? (int)left : NB, b0);
    }
    kept = false;
  }
}

// First pass of the new point (counting) and second pass of the pending point
// (sums) over one traversal.
template <class Add>
struct DualHistVisitor {
  V3f qn, qp;  // new point, pending point
  HistVisitor* hv;
  Sum16CVisitor<Add>* sv;
  __device__ bool far(int a, float split) const {
    const float dn = axis_of(qn, a) - split, dp = axis_of(qp, a) - split;
    const bool same_side = (dn < 0.0f) == (dp < 0.0f);
    return dn * dn <= u2f(hv->bound()) || !same_side || dp * dp <= u2f(sv->bound());
  }
  // a node's box can hold points under either search bound
  __device__ bool near_box(float4 mn, float4 mx) const {
    return box_d2(qn, mn, mx) <= u2f(hv->bound()) || box_d2(qp, mn, mx) <= u2f(sv->bound());
  }
  template <int NB>
  __device__ void points(const float4* p, int cnt, uint32_t base) {
    uint32_t dn[NB], dp[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      dn[j] = f2u(dist2(qn.x, qn.y, qn.z, p[j].x, p[j].y, p[j].z));
      dp[j] = f2u(dist2(qp.x, qp.y, qp.z, p[j].x, p[j].y, p[j].z));
    }
    hv->template points<NB>(dn, cnt);
    sv->template points<NB>(p, dp, cnt, base);
  }
};

// The pending point's second pass on its own (at the end of a march segment,
// or for a point with fewer than k neighbours within the cap): its bound is
// fixed, so it runs on the speculative traversal with node boxes.
template <class Add>
struct SumOnlyVisitor {
  V3f q;
  Sum16CVisitor<Add>* sv;
  __device__ bool far(int a, float split) const {
    const float d = axis_of(q, a) - split;
    return d * d <= u2f(sv->bound());
  }
  __device__ bool near_box(float4 mn, float4 mx) const { return box_d2(q, mn, mx) <= u2f(sv->bound()); }
  template <int NB>
  __device__ void points(const float4* p, int cnt, uint32_t base) {
    uint32_t d[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) d[j] = f2u(dist2(q.x, q.y, q.z, p[j].x, p[j].y, p[j].z));
    sv->template points<NB>(p, d, cnt, base);
  }
};

// A counting first pass on its own (a point gathered alone, seeded with a
// bucket near its k-th neighbour's): the window's bound is fixed, so it runs
// on the speculative traversal with node boxes.
struct HistOnlyVisitor {
  V3f q;
  HistVisitor* hv;
  __device__ bool far(int a, float split) const {
    const float d = axis_of(q, a) - split;
    return d * d <= u2f(hv->bound());
  }
  __device__ bool near_box(float4 mn, float4 mx) const { return box_d2(q, mn, mx) <= u2f(hv->bound()); }
  template <int NB>
  __device__ void points(const float4* p, int cnt, uint32_t) {
    uint32_t d[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) d[j] = f2u(dist2(q.x, q.y, q.z, p[j].x, p[j].y, p[j].z));
    hv->template points<NB>(d, cnt);
  }
};

// The second pass from a seed bucket, counting besides the bucket-Q points
// within the cap (n_eq): with n_less < k <= n_less + n_eq the seed is the
// k-th neighbour's bucket and the sums are those of the second pass from it.
template <class Add>
struct Sum16SeedVisitor : Sum16CVisitor<Add> {
  uint32_t n_eq;
  template <int NB>
  __device__ void points(const float4* p, const uint32_t* dh, int cnt, uint32_t base) {
    if (!this->inclusive) {
#pragma unroll
      for (int j = 0; j < NB; ++j)
        n_eq += j < cnt && dq16(dh[j]) == this->Q && dh[j] <= this->cap_h ? 1u : 0u;
    }
    Sum16CVisitor<Add>::template points<NB>(p, dh, cnt, base);
  }
};

template <class Add>
struct SeedSumVisitor {
  V3f q;
  Sum16SeedVisitor<Add>* sv;
  __device__ bool far(int a, float split) const {
    const float d = axis_of(q, a) - split;
    return d * d <= u2f(sv->bound());
  }
  __device__ bool near_box(float4 mn, float4 mx) const { return box_d2(q, mn, mx) <= u2f(sv->bound()); }
  template <int NB>
  __device__ void points(const float4* p, int cnt, uint32_t base) {
    uint32_t d[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) d[j] = f2u(dist2(q.x, q.y, q.z, p[j].x, p[j].y, p[j].z));
    sv->template points<NB>(p, d, cnt, base);
  }
};

// The heap pass of one point on the speculative traversal with node boxes.
// Its bound shrinks as the heap fills, so a lane decides on far children and
// kept leaves with the bound of that moment, which may be larger than the one
// the leaves are later tested against: more leaves are visited than the
// shrunk bound needs, never fewer.  The heap ends with the k smallest bucket
// keys whatever the order of the points, so Q is that of lane_visit.
struct HeapSpecVisitor {
  V3f q;
  Knn16Visitor* kv;
  __device__ bool far(int a, float split) const {
    const float d = axis_of(q, a) - split;
    return d * d <= u2f(kv->bound());
  }
  __device__ bool near_box(float4 mn, float4 mx) const { return box_d2(q, mn, mx) <= u2f(kv->bound()); }
  template <int NB>
  __device__ void points(const float4* p, int cnt, uint32_t) {
    uint32_t d[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) d[j] = f2u(dist2(q.x, q.y, q.z, p[j].x, p[j].y, p[j].z));
    kv->template points<NB>(p, d, cnt);
  }
};

// SPEC_HEAP: the heap pass on the speculative traversal (HeapSpecVisitor), for
// points that are each gathered on their own; the march's heap passes are rare
// and scattered, and run faster on the plain traversal.
// TOP: the traversals read the top-of-tree tiles the kernel staged in LDS.
template <bool CONE, bool SPEC_HEAP = false, bool TOP = false>
struct LaneGather5 : LaneGather4<CONE, 4> {
  using Base = LaneGather4<CONE, 4>;
  using VolAdd = VolAddT;
  // The window reaches kWinUp buckets above the pending point's bucket Q'
  // (a factor 2^(kWinUp/256) in radius) and 256 - kWinUp below it.
  static constexpr uint32_t kWinUp = 96;
  // The counting window around bucket Qc and its search bound.
  __device__ HistVisitor window(uint32_t Qc, uint32_t caph) const {
    const uint32_t hiq = Qc + kWinUp;
    const uint32_t K0 = hiq > HistVisitor::kBuckets ? hiq - HistVisitor::kBuckets : 0u;
    const uint32_t top = dq16_top(K0 + HistVisitor::kBuckets - 1u);
    return HistVisitor{(lds_u32*)this->hp.H, this->hp.lane, K0, top < caph ? top : caph, 0u, 0u};
  }
  // Q from the counts (full), or the neighbour count when every point within
  // the cap was counted and there are fewer than k; false: take the heap pass.
  __device__ bool resolve(const HistVisitor& hv, uint32_t caph, Knn16Visitor& kv) {
    uint32_t Q;
    if (hv.select((uint32_t)this->gv.k, Q)) {
      kv.full = true;
      kv.root = Q;
      this->hp.size = this->gv.k;
      return true;
    }
    if (hv.top_h == caph && hv.below + hv.inwin < (uint32_t)this->gv.k) {
      this->hp.size = (int)(hv.below + hv.inwin);
      return true;
    }
    return false;
  }

  // Second-pass state of the pending point.
  __device__ Sum16CVisitor<VolAdd> pending_sums(VolAdd& add) const {
    Sum16CVisitor<VolAdd> sv;
    sv.Q = this->pend_Q;
    sv.cap_h = f2u(this->gv.rcap2);
    sv.inclusive = !this->pend_full;
    sv.n_less = 0u;
    sv.d0 = sv.i0 = sv.d1 = sv.i1 = LKEY_INF;
    sv.add = &add;
    return sv;
  }

  // The pending point lies on the current march segment, so its view
  // direction is the wo of the call (it is not kept).
  __device__ bool volume_flush(V3f wo, V3f& c) {
    if (!this->pend) return false;
    this->pend = false;
    if (this->pend_m == 0) {
      c = mul(this->pend_w, v3(0.0f, 0.0f, 0.0f));
      return true;
    }
    VolAdd add{&this->gv, wo, this->hg_g, {}};
    nsum3_init(add.s);
    Sum16CVisitor<VolAdd> sv = pending_sums(add);
    SumOnlyVisitor<VolAdd> w{this->pend_x, &sv};
    lane_visit_dual_spec<4, TOP>(this->gv, this->pend_x.x, this->pend_x.y, this->pend_x.z, w);
    c = this->finish_pending(sv, add);
    return true;
  }

  // The heap pass of x: seeded with bh, then with the cap when fewer than k
  // points fall under the seed (one traversal call site for both).
  __device__ void heap_pass(V3f x, uint32_t bh, uint32_t caph, Knn16Visitor& kv) {
#pragma nounroll
    for (int t = 0; t < 2; ++t) {
      this->hp.size = 0;
      kv.bound_h = t == 0 ? bh : caph;
      if constexpr (SPEC_HEAP) {
        HeapSpecVisitor hs{x, &kv};
        lane_visit_dual_spec<4, TOP>(this->gv, x.x, x.y, x.z, hs);
      } else {
        lane_visit<4>(this->gv, x.x, x.y, x.z, kv);
      }
      if (kv.full || bh == caph) break;
    }
  }

  // The point becomes pending (its view direction is not kept: see
  // volume_flush).
  __device__ void queue(V3f x, V3f w, const Knn16Visitor& kv) {
    this->pend = true;
    this->pend_full = kv.full;
    this->pend_m = this->hp.size;
    this->pend_Q = kv.root;
    this->pend_x = x;
    this->pend_w = w;
  }

  // Seeds are 16-bit buckets (at most 0x7f80), or: no seed, or fewer than k
  // points within the cap.
  static constexpr uint32_t kNoSeed = 0xffffu;
  static constexpr uint32_t kFewSeed = 0xfffeu;

  // One pass from a seed that the query's own previous gather left (the same
  // query on the same photon positions gives the same seed): the second pass
  // from the seed's bucket (or over the cap for kFewSeed), checked against the
  // counts it gathers.  True when the seed was right: kv, hp.size and the sums
  // in sv / add are then those of the first and second passes.  False: sv and
  // add hold nothing usable.
  template <class Add>
  __device__ bool seeded_sums(const GatherMap& gm, V3f x, uint32_t seed, uint32_t caph, Knn16Visitor& kv,
                              Sum16SeedVisitor<Add>& sv, Add& add) {
    const uint32_t k = (uint32_t)gm.k;
    const bool few = seed == kFewSeed;
    if (!few && seed > dq16(caph)) return false;  // not a bucket within the cap
    sv.Q = few ? 0u : seed;
    sv.cap_h = caph;
    sv.inclusive = few;
    sv.n_less = 0u;
    sv.d0 = sv.i0 = sv.d1 = sv.i1 = LKEY_INF;
    sv.add = &add;
    sv.n_eq = 0u;
    SeedSumVisitor<Add> w{x, &sv};
    lane_visit_dual_spec<4, TOP>(gm, x.x, x.y, x.z, w);
    if (few) {
      if (sv.n_less >= k) return false;
      kv.full = false;
      this->hp.size = (int)sv.n_less;
      return true;
    }
    if (sv.n_less >= k || sv.n_less + sv.n_eq < k) return false;
    kv.full = true;
    kv.root = seed;
    this->hp.size = gm.k;
    return true;
  }

  // The first pass of x on its own from a seed bucket Qs near its k-th
  // neighbour's (the same query's bucket in an earlier frame): counting in the
  // window around Qs.  On success kv and hp.size are what the heap pass would
  // leave (Q, or the count when fewer than k points lie within the cap);
  // false when the k-th neighbour lies outside the window or a counter
  // wrapped, and the heap pass must run.  Any seed gives the same result.
  __device__ bool seeded_count(const GatherMap& gm, V3f x, uint32_t Qs, uint32_t caph, Knn16Visitor& kv) {
    HistVisitor hv = window(Qs, caph);
    hv.clear();
    HistOnlyVisitor w{x, &hv};
    lane_visit_dual_spec<4, TOP>(gm, x.x, x.y, x.z, w);
    const uint32_t k = (uint32_t)gm.k;
    uint32_t Q;
    if (hv.select(k, Q)) {
      kv.full = true;
      kv.root = Q;
      this->hp.size = gm.k;
      return true;
    }
    if (hv.top_h == caph && hv.below + hv.inwin < k) {
      this->hp.size = (int)(hv.below + hv.inwin);
      return true;
    }
    return false;
  }

  // The seed for the next gather of the same query: Q, or the cap's bucket
  // when fewer than k points lie within the cap (the window then reaches the
  // cap and counts them all).
  __device__ uint32_t next_seed(const Knn16Visitor& kv) const { return kv.full ? kv.root : kFewSeed; }
  // The counting window for a seed: around the seed's bucket, or reaching the
  // cap for kFewSeed (it then counts every point within the cap).
  __device__ uint32_t window_seed(uint32_t seed, uint32_t caph) const { return seed == kFewSeed ? dq16(caph) : seed; }

  // A surface query on its own, without the cone filter: heap pass (or, from
  // a seed bucket, the counting pass) and second pass on the speculative
  // traversal, then the bucket selection of knn16_gather (pm_lane_knn3.h); the
  // same value as surface() with CONE off, whatever the seed.  *qout (when
  // set) gets the seed for the next gather of the same query.
  __device__ V3f surface_lone(V3f x, V3f n, V3f albedo, uint32_t seed = kNoSeed, uint32_t* qout = nullptr) {
    static_assert(!CONE, "surface_lone has no cone filter");
    ++this->nsurf;
    const GatherMap& gs = this->gs;
    const uint32_t caph = f2u(gs.rcap2);
    this->hp.k = gs.k;
    this->hp.size = 0;
    Knn16Visitor kv{&this->hp, caph, 0u, false};
    SurfAddT add{&gs, n, {}};
    nsum3_init(add.s);
    Sum16SeedVisitor<SurfAddT> sv;
    if (seed == kNoSeed || !seeded_sums(gs, x, seed, caph, kv, sv, add)) {
      if (seed == kNoSeed || !seeded_count(gs, x, window_seed(seed, caph), caph, kv)) {
        this->hp.size = 0;
        HeapSpecVisitor hs{x, &kv};
        lane_visit_dual_spec<4, TOP>(gs, x.x, x.y, x.z, hs);
      }
      if (this->hp.size > 0) {
        nsum3_init(add.s);
        sv.Q = kv.root;
        sv.cap_h = caph;
        sv.inclusive = !kv.full;
        sv.n_less = 0u;
        sv.d0 = sv.i0 = sv.d1 = sv.i1 = LKEY_INF;
        sv.add = &add;
        SumOnlyVisitor<SurfAddT> w{x, &sv};
        lane_visit_dual_spec<4, TOP>(gs, x.x, x.y, x.z, w);
      }
    }
    if (qout) *qout = next_seed(kv);
    if (this->hp.size == 0) return v3(0.0f, 0.0f, 0.0f);
    float r2 = u2f(caph);
    if (kv.full) {
      const uint32_t need = (uint32_t)gs.k - sv.n_less;
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
        Bucket16Visitor bv{sv.Q, kd, ki, LKEY_INF, LKEY_INF};
        lane_visit<4>(gs, x.x, x.y, x.z, bv);
        if (bv.best_i == LKEY_INF) break;
        add(bv.best_i, bv.best_d);
        kd = bv.best_d;
        ki = bv.best_i;
      }
      if (kd != LKEY_INF) r2 = u2f(kd);
    }
    V3f sum = nsum3_get(add.s);
    V3f sc = v3(gs.scale[0], gs.scale[1], gs.scale[2]);
    return mul(mul(sum, albedo), sc) * surface_norm(r2, false);
  }

  // A volume query on its own (nothing pending): the heap pass, or the
  // counting pass from a seed bucket, then the second pass; the value of
  // volume_defer + volume_flush, whatever the seed.  *qout as surface_lone.
  __device__ V3f volume_lone(V3f x, V3f wo, V3f w, uint32_t seed, uint32_t* qout) {
    ++this->nvol;
    const GatherMap& gv = this->gv;
    const uint32_t caph = f2u(gv.rcap2);
    this->hp.k = gv.k;
    this->hp.size = 0;
    Knn16Visitor kv{&this->hp, caph, 0u, false};
    V3f c = v3(0.0f, 0.0f, 0.0f);
    if (seed != kNoSeed) {
      VolAdd add{&gv, wo, this->hg_g, {}};
      nsum3_init(add.s);
      Sum16SeedVisitor<VolAdd> sv;
      if (seeded_sums(gv, x, seed, caph, kv, sv, add)) {
        // the pending point's second pass is done: finish it as volume_flush
        if (qout) *qout = next_seed(kv);
        queue(x, w, kv);
        this->pend = false;
        if (this->pend_m == 0) return mul(this->pend_w, v3(0.0f, 0.0f, 0.0f));
        return this->finish_pending(sv, add);
      }
    }
    if (seed == kNoSeed || !seeded_count(gv, x, window_seed(seed, caph), caph, kv)) {
      this->hp.size = 0;
      this->heap_pass(x, caph, caph, kv);
    }
    if (qout) *qout = next_seed(kv);
    queue(x, w, kv);
    (void)volume_flush(wo, c);
    return c;
  }

  // Surface queries of the march: surface_lone without the cone filter.
  __device__ V3f surface(V3f x, V3f n, V3f albedo, V3f = V3f(), V3f = V3f()) {
    if constexpr (CONE) return Base::surface(x, n, albedo);
    else return surface_lone(x, n, albedo);
  }

  __device__ bool volume_defer(V3f x, V3f wo, V3f w, V3f& c) {
    ++this->nvol;
    const GatherMap& gv = this->gv;
    const uint32_t caph = f2u(gv.rcap2);
    this->hp.k = gv.k;
    this->hp.size = 0;
    Knn16Visitor kv{&this->hp, caph, 0u, false};
    const bool had = this->pend;
    bool heap = true;
    if (had) {
      if (this->pend_m == 0) {
        c = mul(this->pend_w, v3(0.0f, 0.0f, 0.0f));
      } else {
        VolAdd add{&gv, wo, this->hg_g, {}};
        nsum3_init(add.s);
        Sum16CVisitor<VolAdd> sv = pending_sums(add);
        if (this->pend_full) {
          // counting window below and around the pending point's bucket; the
          // search bound is the top of the window (or the cap)
          HistVisitor hv = window(this->pend_Q, caph);
          hv.clear();
          DualHistVisitor<VolAdd> dv{x, this->pend_x, &hv, &sv};
          lane_visit_dual_spec<4, TOP>(gv, x.x, x.y, x.z, dv);
          heap = !resolve(hv, caph, kv);
        } else {
          SumOnlyVisitor<VolAdd> w{this->pend_x, &sv};  // fewer than k within the cap
          lane_visit_dual_spec<4, TOP>(gv, this->pend_x.x, this->pend_x.y, this->pend_x.z, w);
        }
        c = this->finish_pending(sv, add);
      }
    }
    if (heap) {
      // seeded with (r' + |x - x'|)^2 from the last queued point x' when it
      // had k neighbours (r' = top of its bucket): an upper bound of r_k
      uint32_t bh = caph;
      if (this->pend_full) {
        const float pr = sqrt_(u2f(dq16_top(this->pend_Q)));
        float dx = x.x - this->pend_x.x, dy = x.y - this->pend_x.y, dz = x.z - this->pend_x.z;
        float b = (pr + sqrt_((dx * dx + dy * dy) + dz * dz)) * 1.0001f;
        float b2 = b * b;
        if (b2 < gv.rcap2) bh = f2u(b2);
      }
      this->heap_pass(x, bh, caph, kv);
    }
    queue(x, w, kv);
    return had;
  }
};

}  // namespace pm
