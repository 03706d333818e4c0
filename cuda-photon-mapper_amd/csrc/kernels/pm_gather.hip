// pm_gather.hip -- k-nearest-neighbour radiance gather, eye-ray march and
// frame write on gfx950.
//
// Replaces photon_mapping_kernel (photonMappingKernel.cu:1409-1462) with its
// callees computePixelColor (:926-1017), gatherPhotons/integrate (:314-389,
// :586-600) and integrateVolumePhotons (:831-870).  The reference sums a fixed
// 6x6 / 6^3 window of a voxel grid; here each radiance query is an exact kNN
// search over a kd-tree of photon records.
//
// Mapping to CDNA4: one wave64 per query.  The traversal is wave-uniform
// (query, node, stack pointer and pruning bound all live in scalar
// registers), so there is no divergence; the traversal stack lives one entry
// per lane and is read with v_readlane.  A leaf of <= 64 points is one
// coalesced 16-byte-per-lane load.  The current k best (squared distance,
// photon index) keys are kept sorted across the lanes (lane j = j-th
// nearest), which is the register-resident bounded heap: a leaf's candidates
// that beat the current k-th key are found with one ballot, then inserted one
// by one (ballot rank + lane shift) when few, or merged with a 64-lane
// bitonic sort + bitonic merge when many.  Ordering keys are (FP32 bits of
// d^2, photon index), so ties are broken by photon index and the neighbour
// set is exactly the oracle's.
#include "pm_kernels.h"
#include "pm_wave.h"

namespace pm {

static constexpr uint32_t KEY_INF = 0xffffffffu;
static constexpr int SEQ_INSERT_MAX = 12;

// kNN of the query q (identical in all lanes) over map g.  On return lane j <
// count holds the j-th neighbour key (lh = bits(d2), ll = photon index);
// other lanes hold KEY_INF.  Returns count (uniform).
//
// The first n_lds heap-ordered nodes are read from the LDS copy `lds`
// (the top of the tree, shared by the workgroup).  `bound_h` (FP32 bits of a
// squared radius <= rcap^2) only admits keys below (bound_h, KEY_INF): if the
// true k-th distance is within the bound the result is exact; otherwise fewer
// than k neighbours come back and the caller repeats with bound_h = bits(rcap^2).
__device__ __forceinline__ uint32_t wave_knn(const GatherMap& g, const KdNode* lds, uint32_t n_lds, float qx,
                                              float qy, float qz, uint32_t bound_h, uint32_t& lh, uint32_t& ll) {
  const int lane = lane_id();
  lh = KEY_INF;
  ll = KEY_INF;
  const uint32_t n = g.n;
  if (n == 0) return 0u;
  qx = uniform_f(qx);
  qy = uniform_f(qy);
  qz = uniform_f(qz);
  const int k = g.k;
  const uint32_t caph = uniform(bound_h);
  uint32_t wh = caph, wl = KEY_INF;  // current k-th key (pruning bound), uniform
  const uint32_t nl = g.nl;
  // traversal stack: entry i in lane i
  uint32_t s_node = 0u, s_lo = 0u, s_hi = n, s_b = 0u;
  int sp = 1;
  while (sp > 0) {
    --sp;
    uint32_t h = read_lane(s_node, sp);
    uint32_t lo = read_lane(s_lo, sp);
    uint32_t hi = read_lane(s_hi, sp);
    float bound = u2f(read_lane(s_b, sp));
    if (bound > u2f(wh)) continue;
    while (h + 1u < nl) {
      KdNode nd = h < n_lds ? lds[h] : g.nodes[h];
      uint32_t meta = uniform(nd.meta);
      float split = uniform_f(nd.split);
      int a = kd_axis(meta);
      float qa = a == 0 ? qx : (a == 1 ? qy : qz);
      float diff = qa - split;
      uint32_t mid = lo + kd_left_count(meta);
      uint32_t fh, flo, fhi;
      if (diff < 0.0f) {
        fh = 2u * h + 2u; flo = mid; fhi = hi;
        h = 2u * h + 1u; hi = mid;
      } else {
        fh = 2u * h + 1u; flo = lo; fhi = mid;
        h = 2u * h + 2u; lo = mid;
      }
      float fb = diff * diff;
      if (fb <= u2f(wh)) {
        if (lane == sp) {
          s_node = fh; s_lo = flo; s_hi = fhi; s_b = f2u(fb);
        }
        ++sp;
      }
    }
    // leaf [lo, hi), hi - lo <= 64
    uint32_t cnt = hi - lo;
    uint32_t ch = KEY_INF, cl = KEY_INF;
    if ((uint32_t)lane < cnt) {
      float4 p = g.tpts[lo + (uint32_t)lane];
      ch = f2u(dist2(qx, qy, qz, p.x, p.y, p.z));
      cl = f2u(p.w);
    }
    bool pass = key_less(ch, cl, wh, wl);
    uint64_t mask = ballot(pass);
    if (mask == 0ull) continue;
    int c = __popcll(mask);
    if (c <= SEQ_INSERT_MAX) {
      while (mask) {
        int b = __ffsll((unsigned long long)mask) - 1;
        mask &= mask - 1ull;
        uint32_t nh = read_lane(ch, b), nlo = read_lane(cl, b);
        if (!key_less(nh, nlo, wh, wl)) continue;
        int pos = __popcll(ballot(key_less(lh, ll, nh, nlo)));
        uint32_t uh = lane_read(lh, lane > 0 ? lane - 1 : 0);
        uint32_t ul = lane_read(ll, lane > 0 ? lane - 1 : 0);
        if (lane > pos) {
          lh = uh;
          ll = ul;
        } else if (lane == pos) {
          lh = nh;
          ll = nlo;
        }
        if (lane >= k) {
          lh = KEY_INF;
          ll = KEY_INF;
        }
        uint32_t kh = read_lane(lh, k - 1), kl = read_lane(ll, k - 1);
        if (key_less(kh, kl, caph, KEY_INF)) {
          wh = kh;
          wl = kl;
        }
      }
    } else {
      if (!pass) {
        ch = KEY_INF;
        cl = KEY_INF;
      }
      wave_bitonic_sort(ch, cl);
      uint32_t rh = lane_read(ch, 63 - lane), rl = lane_read(cl, 63 - lane);
      if (key_less(rh, rl, lh, ll)) {
        lh = rh;
        ll = rl;
      }
      wave_bitonic_clean(lh, ll);
      if (lane >= k) {
        lh = KEY_INF;
        ll = KEY_INF;
      }
      uint32_t kh = read_lane(lh, k - 1), kl = read_lane(ll, k - 1);
      if (key_less(kh, kl, caph, KEY_INF)) {
        wh = kh;
        wl = kl;
      }
    }
  }
  return (uint32_t)__popcll(ballot(lh != KEY_INF));
}

__device__ __forceinline__ void load_payload(const GatherMap& g, uint32_t idx, V3f& pw, V3f& dir) {
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

// Gather object handed to render_sample (the per-sample eye path shared with
// the oracle).  All lanes of the wave call it with the same arguments.
template <bool CONE>
struct WaveGather {
  GatherMap gs, gv;
  const KdNode* lds_s;
  const KdNode* lds_v;
  uint32_t nlds_s, nlds_v;
  float hg_g;
  uint32_t nsurf, nvol;
  // Previous volume query of this wave: position and k-th neighbour distance.
  // |q - q_prev| + r_prev bounds the new k-th distance (triangle inequality);
  // it only seeds the pruning bound, wave_knn's retry keeps the result exact.
  V3f prev_x;
  float prev_r;  // < 0: none

  __device__ V3f surface(V3f x, V3f n, V3f albedo, V3f = V3f(), V3f = V3f()) {
    ++nsurf;
    uint32_t lh, ll;
    uint32_t m = wave_knn(gs, lds_s, nlds_s, x.x, x.y, x.z, f2u(gs.rcap2), lh, ll);
    if (m == 0u) return v3(0.0f, 0.0f, 0.0f);
    const int lane = lane_id();
    float r2 = (int)m == gs.k ? u2f(read_lane(lh, gs.k - 1)) : gs.rcap2;
    float r = sqrt_(r2);
    V3f v = v3(0.0f, 0.0f, 0.0f);
    if ((uint32_t)lane < m) {
      V3f pw, dir;
      load_payload(gs, ll, pw, dir);
      if (dot(dir, n) < 0.0f) {
        float w = CONE ? cone_weight(u2f(lh), r) : 1.0f;
        v = pw * w;
      }
    }
    V3f sum = v3(wave_sum_comp(v.x), wave_sum_comp(v.y), wave_sum_comp(v.z));
    V3f sc = v3(gs.scale[0], gs.scale[1], gs.scale[2]);
    return mul(mul(sum, albedo), sc) * surface_norm(r2, CONE);
  }

  __device__ V3f volume(V3f x, V3f wo, V3f = V3f()) {
    ++nvol;
    uint32_t lh, ll;
    uint32_t caph = f2u(gv.rcap2);
    uint32_t bh = caph;
    if (prev_r >= 0.0f) {
      float dx = x.x - prev_x.x, dy = x.y - prev_x.y, dz = x.z - prev_x.z;
      float b = (prev_r + sqrt_((dx * dx + dy * dy) + dz * dz)) * 1.0001f;
      float b2 = b * b;
      if (b2 < gv.rcap2) bh = f2u(b2);
    }
    uint32_t m = wave_knn(gv, lds_v, nlds_v, x.x, x.y, x.z, bh, lh, ll);
    if ((int)m < gv.k && bh != caph) m = wave_knn(gv, lds_v, nlds_v, x.x, x.y, x.z, caph, lh, ll);
    prev_x = x;
    prev_r = (int)m == gv.k ? sqrt_(u2f(read_lane(lh, gv.k - 1))) : -1.0f;
    if (m == 0u) return v3(0.0f, 0.0f, 0.0f);
    const int lane = lane_id();
    float r2 = (int)m == gv.k ? u2f(read_lane(lh, gv.k - 1)) : gv.rcap2;
    V3f v = v3(0.0f, 0.0f, 0.0f);
    if ((uint32_t)lane < m) {
      V3f pw, dir;
      load_payload(gv, ll, pw, dir);
      v = pw * phase_hg(dot(dir, wo), hg_g);
    }
    V3f sum = v3(wave_sum_comp(v.x), wave_sum_comp(v.y), wave_sum_comp(v.z));
    V3f sc = v3(gv.scale[0], gv.scale[1], gv.scale[2]);
    return mul(sum, sc) * volume_norm(r2);
  }
};

static constexpr int GATHER_THREADS = 256;
static constexpr int RENDER_THREADS = 512;   // 8 waves share the LDS tree tops
static constexpr uint32_t LDS_NODES = 2047;  // top 11 levels of each map's tree

// One wave per query point; writes k (index, d2) sorted by (d2, index).
__global__ void __launch_bounds__(GATHER_THREADS) k_knn(GatherMap g, const float* __restrict__ q, uint32_t nq,
                                                         uint32_t* __restrict__ idx, float* __restrict__ d2,
                                                         uint32_t* __restrict__ count) {
  uint32_t w = (blockIdx.x * GATHER_THREADS + threadIdx.x) >> 6;
  if (w >= nq) return;
  const int lane = lane_id();
  uint32_t lh, ll;
  uint32_t m = wave_knn(g, nullptr, 0u, q[3 * (size_t)w], q[3 * (size_t)w + 1], q[3 * (size_t)w + 2], f2u(g.rcap2),
                        lh, ll);
  if (lane < g.k) {
    idx[(size_t)w * g.k + lane] = ll;
    d2[(size_t)w * g.k + lane] = u2f(lh == KEY_INF ? 0x7f800000u : lh);
  }
  if (lane == 0 && count) count[w] = m;
}

// Persistent render: a fixed grid sized to the GPU; wave i renders pixels
// i, i + W, i + 2W, ... of the crop (W = waves in the grid), looping over each
// pixel's samples.  Each workgroup first copies the top LDS_NODES nodes of
// both trees into LDS.
template <bool FP64T, bool CONE>
__global__ void __launch_bounds__(RENDER_THREADS) k_render(RenderLaunch r) {
  __shared__ KdNode top_s[LDS_NODES];
  __shared__ KdNode top_v[LDS_NODES];
  const uint32_t ns = r.gs.nl > 1u ? min(LDS_NODES, r.gs.nl - 1u) : 0u;
  const uint32_t nv = r.gv.nl > 1u ? min(LDS_NODES, r.gv.nl - 1u) : 0u;
  for (uint32_t i = threadIdx.x; i < ns; i += RENDER_THREADS) top_s[i] = r.gs.nodes[i];
  for (uint32_t i = threadIdx.x; i < nv; i += RENDER_THREADS) top_v[i] = r.gv.nodes[i];
  __syncthreads();
  const int lane = lane_id();
  const uint32_t waves = gridDim.x * (RENDER_THREADS / 64);
  const uint32_t npix = (uint32_t)(r.w * r.h);
  WaveGather<CONE> g;
  g.gs = r.gs;
  g.gv = r.gv;
  g.lds_s = top_s;
  g.lds_v = top_v;
  g.nlds_s = ns;
  g.nlds_v = nv;
  g.hg_g = r.sc.med.g;
  g.nsurf = 0u;
  g.nvol = 0u;
  g.prev_r = -1.0f;
  for (uint32_t w = uniform(blockIdx.x * (RENDER_THREADS / 64) + (threadIdx.x >> 6)); w < npix; w += waves) {
    uint32_t px = (uint32_t)r.x0 + w % (uint32_t)r.w;
    uint32_t py = (uint32_t)r.y0 + w / (uint32_t)r.w;
    NSum3<float> acc;
    nsum3_init(acc);
    for (int s = 0; s < r.rp.spp; ++s) {
      V3f v = FP64T ? render_sample_t<float, double>(r.sc, r.rp, px, py, (uint32_t)s, g)
                    : render_sample_t<float, float>(r.sc, r.rp, px, py, (uint32_t)s, g);
      nsum3_add(acc, v);
    }
    V3f v = nsum3_get(acc) * (1.0f / (float)r.rp.spp);
    if (lane == 0) {
      if (r.rgb) {
        r.rgb[3 * (size_t)w] = v.x;
        r.rgb[3 * (size_t)w + 1] = v.y;
        r.rgb[3 * (size_t)w + 2] = v.z;
      }
      if (r.u8) {
        uchar4 o;
        o.x = tonemap_u8(v.x, r.exposure);
        o.y = tonemap_u8(v.y, r.exposure);
        o.z = tonemap_u8(v.z, r.exposure);
        o.w = 0;
        r.u8[w] = o;
      }
    }
  }
  if (lane == 0 && r.counters) {
    atomicAdd(&r.counters[0], (unsigned long long)g.nsurf);
    atomicAdd(&r.counters[1], (unsigned long long)g.nvol);
  }
}

hipError_t launch_knn(GatherMap g, const float* q, uint32_t nq, uint32_t* idx, float* d2, uint32_t* count,
                      hipStream_t st) {
  if (nq == 0) return hipSuccess;
  uint64_t threads = (uint64_t)nq * 64u;
  unsigned nb = (unsigned)((threads + GATHER_THREADS - 1) / GATHER_THREADS);
  hipLaunchKernelGGL(k_knn, dim3(nb), dim3(GATHER_THREADS), 0, st, g, q, nq, idx, d2, count);
  return hipGetLastError();
}

hipError_t launch_render(const RenderLaunch& r, hipStream_t st) {
  uint64_t npix = (uint64_t)r.w * (uint64_t)r.h;
  if (npix == 0) return hipSuccess;
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  uint64_t want = (npix + RENDER_THREADS / 64 - 1) / (RENDER_THREADS / 64);
  uint64_t nb = (uint64_t)(cus > 0 ? cus : 1) * (uint64_t)r.blocks_per_cu;
  if (nb > want) nb = want;
  const bool fp64 = r.rp.fp64_transmittance != 0, cone = r.estimator == 1;
  if (fp64 && cone) hipLaunchKernelGGL((k_render<true, true>), dim3((unsigned)nb), dim3(RENDER_THREADS), 0, st, r);
  else if (fp64) hipLaunchKernelGGL((k_render<true, false>), dim3((unsigned)nb), dim3(RENDER_THREADS), 0, st, r);
  else if (cone) hipLaunchKernelGGL((k_render<false, true>), dim3((unsigned)nb), dim3(RENDER_THREADS), 0, st, r);
  else hipLaunchKernelGGL((k_render<false, false>), dim3((unsigned)nb), dim3(RENDER_THREADS), 0, st, r);
  return hipGetLastError();
}

}  // namespace pm
