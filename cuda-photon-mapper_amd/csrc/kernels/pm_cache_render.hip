// pm_cache_render.hip -- render with the learned radiance cache on gfx950.
//
// Same eye paths as the plain render (render_sample, shared with the oracle),
// but every radiance query is first recorded instead of gathered:
//   1. k_rec (count)  : per pixel, the number of queries of its samples
//   2. scan           : per-pixel record offsets (records stay in pixel order)
//   3. k_rec (write)  : query inputs (position, direction, normal or 0),
//                       path weights (throughput x transmittance x step, with
//                       the surface albedo folded in) and record codes (error
//                       cell, kind, near-light bit: record_code)
//   4. training       : a Philox-chosen subset of the records gets exact kNN
//                       targets (k_values7 under gather modes 7 and auto,
//                       k_values otherwise) and trains the MLP (pm_mlp.hip);
//                       each batch's pre-update errors are binned into a
//                       16^3 x {surface, volume} grid (by the fused forward
//                       pass, k_fwd_fused in pm_mlp_train.hip)
//   5. inference      : MLP over the volume records (all records when surface
//                       records are cached too; pm_mlp_infer.hip)
//   6. fallback       : records in cells whose mean training error exceeds
//                       the threshold (or that saw no training sample) are
//                       re-gathered exactly with kNN (k_cell_fb, the lists of
//                       record_class by k_list_count / k_list_scatter, then
//                       k_values7 / k_values over the surface rows and the
//                       volume rows)
//   7. k_accumulate   : per pixel, compensated sum of weight x value / spp.
// No reference counterpart (the reference has no cache); the plain path it
// replaces is photon_mapping_kernel, photonMappingKernel.cu:1409-1462.
#include "pm_lane_knn5.h"
#include "pm_mlp.h"

namespace pm {

// A record's code, everything the fallback rule reads of it: its error cell
// (bits 0-12; bit 12 set for volume records) and, in bit 13, whether a volume
// record lies within the near-light radius.
constexpr uint32_t kCodeVolume = (uint32_t)(CACHE_CELLS / 2), kCodeNear = (uint32_t)CACHE_CELLS;
static_assert(CACHE_CELLS == 8192, "record codes hold 13 cell bits and the near-light bit");
__device__ __forceinline__ uint16_t record_code(V3f x, V3f nrm, V3f light, float near2) {
  const bool surf = !(nrm.x == 0.0f && nrm.y == 0.0f && nrm.z == 0.0f);
  const float lx = x.x - light.x, ly = x.y - light.y, lz = x.z - light.z;
  const bool near = !surf && (lx * lx + ly * ly) + lz * lz < near2;
  return (uint16_t)(err_cell_of(x.x, x.y, x.z, surf) | (near ? kCodeNear : 0u));
}

// Recording gather: stores each query and its weight, returns zero radiance.
// With march stride s > 1, one volume record stands for up to s consecutive
// march points of the same segment: its weight is the sum of theirs and its
// position their weight-weighted centroid (a midpoint rule along the ray,
// second order in the step, where the first point alone would be first
// order).  A point farther than 1.5 s steps from the open record's first
// point, a surface query, or a point on a ray of another direction starts a
// new record.  Within the near-light radius (near2) of the point light,
// where the in-scattered radiance grows as 1 / distance^2 and no low-order
// rule over s steps holds, every march point keeps its own record (and is
// gathered exactly, see record_class).  The open volume record stays in registers
// and is written once, when the next record starts or the pixel ends (flush).
struct RecordGather {
  bool write;
  uint32_t n;        // records of this lane so far
  uint32_t base;     // first record index of this pixel (write pass)
  uint32_t pixel;    // crop pixel index
  float* in;
  float* wt;
  uint16_t* code;
  int stride;
  float reach2;      // (1.5 * stride * step)^2
  bool pending;      // the open volume record is not written yet
  uint32_t open_i;   // its index among this lane's records
  int open;          // points merged into the open volume record (0: it takes no more)
  V3f open_x, open_d;
  V3f open_w;        // sum of the merged points' weights
  float open_sw;     // sum of the merged points' scalar weights (w.x + w.y + w.z)
  V3f open_swx;      // sum of scalar weight x position
  V3f light;
  float near2;       // squared near-light radius: no merging closer than this to the light
  __device__ void store(uint32_t k, V3f x, V3f d, V3f nrm, V3f w) {
    if (!write) return;
    size_t i = (size_t)base + k;
    float* p = in + i * MLP_IN;
    p[0] = x.x; p[1] = x.y; p[2] = x.z;
    p[3] = d.x; p[4] = d.y; p[5] = d.z;
    p[6] = nrm.x; p[7] = nrm.y; p[8] = nrm.z;
    wt[3 * i] = w.x; wt[3 * i + 1] = w.y; wt[3 * i + 2] = w.z;
    code[i] = record_code(x, nrm, light, near2);
  }
  // Writes the open volume record: at the weighted centroid of its points
  // (the first point while there is one, or no weight).
  __device__ void flush() {
    if (!pending) return;
    pending = false;
    V3f p = open_x;
    if (open > 1 && open_sw > 0.0f) {
      const float inv = 1.0f / open_sw;
      p = v3(open_swx.x * inv, open_swx.y * inv, open_swx.z * inv);
    }
    store(open_i, p, open_d, v3(0.0f, 0.0f, 0.0f), open_w);
  }
  __device__ V3f surface(V3f x, V3f nrm, V3f albedo, V3f w, V3f wo) {
    flush();
    open = 0;
    store(n++, x, wo, nrm, mul(w, albedo));
    return v3(0.0f, 0.0f, 0.0f);
  }
  __device__ V3f volume(V3f x, V3f wo, V3f w) {
    const float lx = x.x - light.x, ly = x.y - light.y, lz = x.z - light.z;
    const bool near = (lx * lx + ly * ly) + lz * lz < near2;
    if (!near && open > 0 && open < stride && wo.x == open_d.x && wo.y == open_d.y && wo.z == open_d.z) {
      float dx = x.x - open_x.x, dy = x.y - open_x.y, dz = x.z - open_x.z;
      if ((dx * dx + dy * dy) + dz * dz <= reach2) {
        const float sw = (w.x + w.y) + w.z;
        open_sw = open_sw + sw;
        open_swx = open_swx + x * sw;
        open_w = open_w + w;
        ++open;
        return v3(0.0f, 0.0f, 0.0f);
      }
    }
    flush();
    pending = true;
    open_i = n++;
    open = near ? 0 : 1;
    open_x = x;
    open_d = wo;
    open_w = w;
    open_sw = (w.x + w.y) + w.z;
    open_swx = x * open_sw;
    return v3(0.0f, 0.0f, 0.0f);
  }
};

// The counting pass: the number of records of a path does not depend on its
// weights, so the march skips the transmittance.
struct RecordCount : RecordGather {
  static constexpr bool kCountOnly = true;
};

// One lane per crop pixel, 8x8 tiles per wave (same pixel order as the lane
// render); counts and offsets by record-order slot.  COUNT: the counting pass
// (r.write == 0), which also clears the padding slots of partial blocks.
template <bool FP64T, bool COUNT>
__global__ void __launch_bounds__(64) k_rec(CacheRecordLaunch r) {
  const int lane = lane_id();
  const int tiles_x = (r.w + 7) / 8;
  const int tile = (int)blockIdx.x;
  const int lx = (tile % tiles_x) * 8 + (lane & 7), ly = (tile / tiles_x) * 8 + (lane >> 3);
  if (lx >= r.w || ly >= r.h) {
    if (!r.write && lx < (r.w + 3) / 4 * 4 && ly < (r.h + 3) / 4 * 4) r.counts[cache_pixel_order(lx, ly, r.w)] = 0u;
    return;
  }
  const uint32_t pidx = cache_pixel_order(lx, ly, r.w);
  typename std::conditional<COUNT, RecordCount, RecordGather>::type g;
  g.write = r.write != 0;
  g.n = 0;
  g.base = r.write ? r.offsets[pidx] : 0u;
  g.pixel = pidx;
  g.in = r.in;
  g.wt = r.wt;
  g.code = r.code;
  g.stride = r.stride > 0 ? r.stride : 1;
  const float reach = 1.5f * (float)g.stride * r.rp.march_step;
  g.reach2 = reach * reach;
  g.pending = false;
  g.open = 0;
  g.light = v3(r.sc.light[0], r.sc.light[1], r.sc.light[2]);
  g.near2 = r.near2;
  const uint32_t px = (uint32_t)(r.x0 + lx), py = (uint32_t)(r.y0 + ly);
  for (int s = 0; s < r.rp.spp; ++s) {
    if (FP64T) (void)render_sample_t<float, double>(r.sc, r.rp, px, py, (uint32_t)s, g);
    else (void)render_sample_t<float, float>(r.sc, r.rp, px, py, (uint32_t)s, g);
  }
  g.flush();
  if (!r.write) r.counts[pidx] = g.n;
}

// Exact kNN value (albedo-free for surfaces) of records list[i] (or i).
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) k_values(CacheValueLaunch v) {
  extern __shared__ uint32_t heap_lds[];
  const int lane = lane_id();
  const uint32_t i = blockIdx.x * 64u + (uint32_t)lane;
  if (i >= v.n) return;
  const uint32_t rec = v.list ? v.list[i] : i;
  const float* p = v.in + (size_t)rec * MLP_IN;
  LaneGather3<false, 4> g;  // 2-byte heap: 8 KB of LDS per wave
  g.gs = v.gs;
  g.gv = v.gv;
  g.hp.H = (uint16_t*)heap_lds;
  g.hp.lane = lane;
  g.hp.k = 0;
  g.hp.size = 0;
  g.hg_g = v.hg_g;
  g.nsurf = g.nvol = 0u;
  g.prev_r = -1.0f;
  V3f x = v3(p[0], p[1], p[2]), d = v3(p[3], p[4], p[5]), nrm = v3(p[6], p[7], p[8]);
  const bool surf = !(nrm.x == 0.0f && nrm.y == 0.0f && nrm.z == 0.0f);
  V3f val = surf ? g.surface(x, nrm, v3(1.0f, 1.0f, 1.0f), v3(1.0f, 1.0f, 1.0f), d) : g.volume(x, d, v3(1.0f, 1.0f, 1.0f));
  const size_t o = v.out_by_list ? (size_t)rec : (size_t)i;
  v.out[3 * o] = val.x;
  v.out[3 * o + 1] = val.y;
  v.out[3 * o + 2] = val.z;
}

__device__ __forceinline__ void put_value(const CacheValueLaunch& v, uint32_t o, V3f val) {
  v.out[3 * (size_t)o] = val.x;
  v.out[3 * (size_t)o + 1] = val.y;
  v.out[3 * (size_t)o + 2] = val.z;
}

// The same values with the gather of mode 7 (pm_lane_knn5.h), one row per
// lane: a heap pass for Q on the speculative traversal with node boxes, then
// the second pass (also speculative, from the tree-ordered payload), for
// surface and volume rows alike.  A row has no previous march point to seed a
// count window from.  Values equal k_values'.  KIND 1 or 2: every row is a
// surface or a volume query (the other branch is not compiled in).
template <int KIND>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) k_values7(CacheValueLaunch v) {
  extern __shared__ uint32_t heap_lds[];
  const int lane = lane_id();
  const uint32_t i = blockIdx.x * 64u + (uint32_t)lane;
  if (i >= v.n) return;
  const uint32_t rec = v.list ? v.list[i] : i;
  const uint32_t o = v.out_by_list ? rec : i;
  const float* p = v.in + (size_t)rec * MLP_IN;
  LaneGather5<false, true> g;
  g.gs = v.gs;
  g.gv = v.gv;
  g.hp.H = (uint16_t*)heap_lds;
  g.hp.lane = lane;
  g.hp.k = 0;
  g.hp.size = 0;
  g.hg_g = v.hg_g;
  g.nsurf = g.nvol = 0u;
  g.prev_r = -1.0f;
  g.pend = false;
  g.pend_full = false;
  const V3f one = v3(1.0f, 1.0f, 1.0f);
  V3f x = v3(p[0], p[1], p[2]), d = v3(p[3], p[4], p[5]), nrm = v3(p[6], p[7], p[8]);
  V3f c;
  const bool surf = KIND == 0 ? !(nrm.x == 0.0f && nrm.y == 0.0f && nrm.z == 0.0f) : KIND == 1;
  // the seed bucket of this record from the previous frame (none: heap pass)
  using G = LaneGather5<false, true>;
  const uint32_t seed = v.seed_q ? (uint32_t)v.seed_q[rec] : G::kNoSeed;
  uint32_t q = G::kNoSeed;
  if (surf) {
    c = g.surface_lone(x, nrm, one, seed, &q);
  } else {
    c = g.volume_lone(x, d, one, seed, &q);
  }
  if (v.seed_q) v.seed_q[rec] = (uint16_t)q;
  put_value(v, o, c);
}

// Training subset, stratified by query kind: rows j < count/2 are drawn from
// the surface records (list_s, n_s of them), the rest from the volume records
// (list_v, n_v), each uniformly with a Philox draw (surface queries are few
// per pixel but carry most of the image).  An empty kind gives its share to
// the other.
__global__ void k_select(const float* __restrict__ in, const uint32_t* __restrict__ list_s, uint32_t n_s,
                         const uint32_t* __restrict__ list_v, uint32_t n_v, uint32_t count, uint32_t seed,
                         uint32_t frame, float* __restrict__ out) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  U4 r = rng(seed, j, frame, 0u, TAG_CACHE);
  uint64_t u = ((uint64_t)r.x << 32) | r.y;
  bool surf = (j < count / 2u && n_s > 0u) || n_v == 0u;
  const uint32_t* list = surf ? list_s : list_v;
  uint64_t n = surf ? n_s : n_v;
  uint64_t idx = list[(uint64_t)(((unsigned __int128)u * n) >> 64)];
  for (int c = 0; c < MLP_IN; ++c) out[(size_t)j * MLP_IN + c] = in[idx * MLP_IN + c];
}

__global__ void k_row_codes(const float* __restrict__ in, uint64_t n, uint16_t* __restrict__ code) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) code[i] = (uint16_t)err_cell(in + i * MLP_IN);
}


// Fallback decision per error cell: no training row in the cell, or a mean
// (over its training rows and channels) squared log error above the threshold.
__global__ void k_cell_fb(const unsigned long long* __restrict__ esum, const unsigned int* __restrict__ ecnt,
                          float threshold, uint8_t* __restrict__ cell_fb) {
  const uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= (uint32_t)CACHE_CELLS) return;
  const uint32_t c = ecnt[cell];
  const double mse = (double)esum[cell] / 16777216.0 / ((double)c * (double)MLP_OUT);
  cell_fb[cell] = c == 0u || mse > (double)threshold ? 1u : 0u;
}

// The list (0 or 1) a record goes into, or -1.  Fallback lists: surface
// records (unless surface records are cached), volume records within the
// near-light radius (recorded one per march point, where the radiance is
// steepest), and the records of cells marked by k_cell_fb.
__device__ __forceinline__ int record_class(uint32_t code, const CacheListRule& r) {
  const bool vol = (code & kCodeVolume) != 0u;
  if (!r.fallback) return vol ? 1 : 0;
  const bool cfb = r.cell_fb[code & (kCodeNear - 1u)] != 0;
  if (!vol) return !r.cache_surface || cfb ? 0 : -1;
  return (code & kCodeNear) || cfb ? 1 : -1;
}

// List building: a block takes LIST_TILE consecutive records, LIST_ITEMS
// consecutive ones per thread (two 16-byte code loads).
constexpr int LIST_THREADS = 256, LIST_ITEMS = 16, LIST_TILE = LIST_THREADS * LIST_ITEMS;

__device__ __forceinline__ void load_codes(const uint16_t* __restrict__ code, uint64_t n, uint64_t base,
                                           uint16_t (&v)[LIST_ITEMS]) {
  if (base + LIST_ITEMS <= n) {
    const uint4 a = *(const uint4*)&code[base], b = *(const uint4*)&code[base + 8];
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      v[2 * r] = (uint16_t)(w[r] & 0xffffu);
      v[2 * r + 1] = (uint16_t)(w[r] >> 16);
    }
  } else {
#pragma unroll
    for (int r = 0; r < LIST_ITEMS; ++r) v[r] = base + r < n ? code[base + r] : (uint16_t)0;
  }
}

// The two lists' counts of a thread's records, packed (list 0 in the low 16
// bits, list 1 in the high 16; a block's totals are at most LIST_TILE each).
__device__ __forceinline__ uint32_t class_counts(const uint16_t (&v)[LIST_ITEMS], uint64_t n, uint64_t base,
                                                 const CacheListRule& rule) {
  uint32_t c = 0u;
#pragma unroll
  for (int r = 0; r < LIST_ITEMS; ++r) {
    const int k = base + r < n ? record_class(v[r], rule) : -1;
    c += k == 0 ? 1u : (k == 1 ? 0x10000u : 0u);
  }
  return c;
}

// Exclusive scan of one packed value per thread over the block; *total gets
// the block's sum.
__device__ __forceinline__ uint32_t list_block_scan(uint32_t v, uint32_t* lds_waves, uint32_t* total) {
  const int lane = lane_id(), wave = (int)(threadIdx.x >> 6);
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = lane_read(x, lane >= d ? lane - d : lane);
    if (lane >= d) x += y;
  }
  if (lane == 63) lds_waves[wave] = x;
  __syncthreads();
  uint32_t before = 0u, tot = 0u;
#pragma unroll
  for (int w = 0; w < LIST_THREADS / 64; ++w) {
    const uint32_t s = lds_waves[w];
    before += w < wave ? s : 0u;
    tot += s;
  }
  *total = tot;
  return before + x - v;
}

__global__ void __launch_bounds__(LIST_THREADS) k_list_count(const uint16_t* __restrict__ code, uint64_t n,
                                                             CacheListRule rule, uint32_t* __restrict__ counts) {
  __shared__ uint32_t lw[LIST_THREADS / 64];
  const uint64_t base = (uint64_t)blockIdx.x * LIST_TILE + (uint64_t)threadIdx.x * LIST_ITEMS;
  uint16_t v[LIST_ITEMS];
  load_codes(code, n, base, v);
  uint32_t tot;
  (void)list_block_scan(class_counts(v, n, base, rule), lw, &tot);
  if (threadIdx.x == 0) {
    counts[blockIdx.x] = tot & 0xffffu;
    counts[gridDim.x + blockIdx.x] = tot >> 16;
  }
}

// The block's entries of both lists are staged in LDS (list 0's, then list
// 1's: at most LIST_TILE in all) and written out as two contiguous runs.
__global__ void __launch_bounds__(LIST_THREADS) k_list_scatter(const uint16_t* __restrict__ code, uint64_t n,
                                                               CacheListRule rule, const uint32_t* __restrict__ offs,
                                                               uint32_t* __restrict__ list) {
  __shared__ uint32_t lw[LIST_THREADS / 64];
  __shared__ uint32_t stage[LIST_TILE];
  const uint64_t tile0 = (uint64_t)blockIdx.x * LIST_TILE;
  const uint64_t base = tile0 + (uint64_t)threadIdx.x * LIST_ITEMS;
  uint16_t v[LIST_ITEMS];
  load_codes(code, n, base, v);
  uint32_t tot;
  const uint32_t ex = list_block_scan(class_counts(v, n, base, rule), lw, &tot);
  const uint32_t n0 = tot & 0xffffu, n1 = tot >> 16;
  uint32_t l0 = ex & 0xffffu, l1 = n0 + (ex >> 16);
#pragma unroll
  for (int r = 0; r < LIST_ITEMS; ++r) {
    const int k = base + r < n ? record_class(v[r], rule) : -1;
    const uint32_t local = (uint32_t)(threadIdx.x * LIST_ITEMS + r);  // < LIST_TILE: fits the record index
    if (k == 0) stage[l0++] = local;
    if (k == 1) stage[l1++] = local;
  }
  __syncthreads();
  const uint32_t o0 = offs[blockIdx.x], o1 = offs[gridDim.x + blockIdx.x];
  for (uint32_t i = threadIdx.x; i < n0; i += LIST_THREADS) list[o0 + i] = (uint32_t)(tile0 + stage[i]);
  for (uint32_t i = threadIdx.x; i < n1; i += LIST_THREADS) list[o1 + i] = (uint32_t)(tile0 + stage[n0 + i]);
}

// acc += sum of v[0..n) in a fixed order: workgroup b sums its chunk of
// SUM_CHUNK values into part[b] (a fixed tree), then one workgroup adds the
// partials in order.  The frame's training loss without a host round trip
// per step.
constexpr uint32_t SUM_CHUNK = 4096, SUM_PARTS = 256;
__global__ void __launch_bounds__(256) k_sum_rows(const float* __restrict__ v, uint32_t n, double* __restrict__ part) {
  __shared__ double red[256];
  const uint32_t b0 = blockIdx.x * SUM_CHUNK, b1 = min(n, b0 + SUM_CHUNK);
  double s = 0.0;
  for (uint32_t i = b0 + threadIdx.x; i < b1; i += 256u) s += (double)v[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_sum_parts(const double* __restrict__ part, uint32_t np, double* __restrict__ acc) {
  double s = 0.0;
  for (uint32_t b = 0; b < np; ++b) s += part[b];
  *acc += s;
}

// Per pixel: compensated sum over its records of weight * value, / spp.
__global__ void k_accumulate(const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt, int w, int h,
                             const float* __restrict__ wt, const float* __restrict__ val, int spp, float exposure,
                             float* __restrict__ rgb, uchar4* __restrict__ u8) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (uint32_t)w * (uint32_t)h) return;
  const uint32_t o = cache_pixel_order((int)(p % (uint32_t)w), (int)(p / (uint32_t)w), w);
  NSum3<float> acc;
  nsum3_init(acc);
  const size_t b = off[o], e = b + cnt[o];
  size_t i = b;
  // four records' loads issued before their adds (the adds stay in record order)
  for (; i + 4 <= e; i += 4) {
    V3f w[4], x[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const size_t r = i + (size_t)j;
      w[j] = v3(wt[3 * r], wt[3 * r + 1], wt[3 * r + 2]);
      x[j] = v3(val[3 * r], val[3 * r + 1], val[3 * r + 2]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) nsum3_add(acc, mul(w[j], x[j]));
  }
  for (; i < e; ++i)
    nsum3_add(acc, mul(v3(wt[3 * i], wt[3 * i + 1], wt[3 * i + 2]), v3(val[3 * i], val[3 * i + 1], val[3 * i + 2])));
  V3f v = nsum3_get(acc) * (1.0f / (float)spp);
  if (rgb) {
    rgb[3 * (size_t)p] = v.x;
    rgb[3 * (size_t)p + 1] = v.y;
    rgb[3 * (size_t)p + 2] = v.z;
  }
  if (u8) {
    uchar4 q;
    q.x = tonemap_u8(v.x, exposure);
    q.y = tonemap_u8(v.y, exposure);
    q.z = tonemap_u8(v.z, exposure);
    q.w = 0;
    u8[p] = q;
  }
}

hipError_t launch_cache_record(const CacheRecordLaunch& r, hipStream_t st) {
  unsigned tiles = (unsigned)(((r.w + 7) / 8) * ((r.h + 7) / 8));
  if (tiles == 0) return hipSuccess;
  if (!r.write) hipLaunchKernelGGL((k_rec<false, true>), dim3(tiles), dim3(64), 0, st, r);
  else if (r.rp.fp64_transmittance) hipLaunchKernelGGL((k_rec<true, false>), dim3(tiles), dim3(64), 0, st, r);
  else hipLaunchKernelGGL((k_rec<false, false>), dim3(tiles), dim3(64), 0, st, r);
  return hipGetLastError();
}

hipError_t launch_cache_values(const CacheValueLaunch& v, hipStream_t st) {
  if (v.n == 0) return hipSuccess;
  const size_t lds = 64 * 64 * sizeof(uint16_t);  // 2-byte heap of 64 keys per lane
  const dim3 grid((v.n + 63) / 64);
  if (!v.mode7) hipLaunchKernelGGL(k_values, grid, dim3(64), lds, st, v);
  else if (v.kind == 1) hipLaunchKernelGGL(k_values7<1>, grid, dim3(64), lds, st, v);
  else if (v.kind == 2) hipLaunchKernelGGL(k_values7<2>, grid, dim3(64), lds, st, v);
  else hipLaunchKernelGGL(k_values7<0>, grid, dim3(64), lds, st, v);
  return hipGetLastError();
}

hipError_t launch_cache_select(const float* in, const uint32_t* list_s, uint32_t n_s, const uint32_t* list_v,
                               uint32_t n_v, uint32_t count, uint32_t seed, uint32_t frame, float* out, hipStream_t st) {
  if (count == 0 || (n_s == 0 && n_v == 0)) return hipSuccess;
  hipLaunchKernelGGL(k_select, dim3((count + 255) / 256), dim3(256), 0, st, in, list_s, n_s, list_v, n_v, count, seed,
                     frame, out);
  return hipGetLastError();
}


// *total += the sum of the n per-pixel record counts, in 64 bits: one 64-bit
// atomic add per workgroup (integer sums: any order gives the same total).
__global__ void __launch_bounds__(256) k_count_total(const uint32_t* __restrict__ cnt, uint32_t n,
                                                     unsigned long long* __restrict__ total) {
  __shared__ unsigned long long part[256 / 64];
  unsigned long long s = 0ull;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) s += cnt[i];
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) s += __shfl_xor(s, d);
  if ((threadIdx.x & 63u) == 0u) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(total, part[0] + part[1] + part[2] + part[3]);
}

__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

// *out += a 64-bit fingerprint of n points' positions (the wrapping sum of a
// mix of each point's coordinates and position in the array): equal for equal
// arrays, in any summation order.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__global__ void __launch_bounds__(256) k_point_hash(const float4* __restrict__ p, uint32_t n,
                                                    unsigned long long* __restrict__ out) {
  __shared__ unsigned long long part[256 / 64];
  unsigned long long s = 0ull;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    const float4 q = p[i];
    const unsigned long long a = ((unsigned long long)f2u(q.x) << 32) | f2u(q.y);
    s += mix64(a ^ mix64(((unsigned long long)f2u(q.z) << 32) | i));
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) s += __shfl_xor(s, d);
  if ((threadIdx.x & 63u) == 0u) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, part[0] + part[1] + part[2] + part[3]);
}

hipError_t launch_point_hash(const float4* p, uint32_t n, unsigned long long* out, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, out);
  if (n == 0) return hipGetLastError();
  const uint32_t blocks = std::max(1u, std::min(1024u, (n + 2047u) / 2048u));
  hipLaunchKernelGGL(k_point_hash, dim3(blocks), dim3(256), 0, st, p, n, out);
  return hipGetLastError();
}

hipError_t launch_count_total(const uint32_t* cnt, uint32_t n, unsigned long long* total, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, total);
  const uint32_t blocks = std::max(1u, std::min(1024u, (n + 2047u) / 2048u));
  hipLaunchKernelGGL(k_count_total, dim3(blocks), dim3(256), 0, st, cnt, n, total);
  return hipGetLastError();
}

size_t sum_rows_parts() { return SUM_PARTS; }

hipError_t launch_sum_rows(const float* v, uint32_t n, double* acc, double* part, hipStream_t st) {
  if (n == 0) return hipSuccess;
  for (uint32_t o = 0; o < n; o += SUM_CHUNK * SUM_PARTS) {
    const uint32_t m = std::min(n - o, SUM_CHUNK * SUM_PARTS);
    const uint32_t np = (m + SUM_CHUNK - 1) / SUM_CHUNK;
    hipLaunchKernelGGL(k_sum_rows, dim3(np), dim3(256), 0, st, v + o, m, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(1), 0, st, part, np, acc);
  }
  return hipGetLastError();
}

hipError_t launch_cache_cell_fb(const unsigned long long* esum, const unsigned int* ecnt, float threshold,
                                uint8_t* cell_fb, hipStream_t st) {
  hipLaunchKernelGGL(k_cell_fb, dim3((CACHE_CELLS + 255) / 256), dim3(256), 0, st, esum, ecnt, threshold, cell_fb);
  return hipGetLastError();
}

size_t cache_list_blocks(uint64_t n) { return (size_t)((n + LIST_TILE - 1) / LIST_TILE); }

hipError_t launch_cache_row_codes(const float* in, uint64_t n, uint16_t* code, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_row_codes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, n, code);
  return hipGetLastError();
}

hipError_t launch_cache_list_count(const uint16_t* code, uint64_t n, const CacheListRule& rule, uint32_t* counts,
                                   hipStream_t st) {
  if (n == 0) return hipSuccess;
  if (rule.fallback && !rule.cell_fb) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_list_count, dim3((unsigned)cache_list_blocks(n)), dim3(LIST_THREADS), 0, st, code, n, rule,
                     counts);
  return hipGetLastError();
}

hipError_t launch_cache_list_scatter(const uint16_t* code, uint64_t n, const CacheListRule& rule,
                                     const uint32_t* offs, uint32_t* list, hipStream_t st) {
  if (n == 0) return hipSuccess;
  if (rule.fallback && !rule.cell_fb) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_list_scatter, dim3((unsigned)cache_list_blocks(n)), dim3(LIST_THREADS), 0, st, code, n, rule,
                     offs, list);
  return hipGetLastError();
}

hipError_t launch_cache_accumulate(const uint32_t* off, const uint32_t* cnt, int w, int h, const float* wt,
                                   const float* val, int spp, float exposure, float* rgb, uchar4* u8, hipStream_t st) {
  const uint32_t npix = (uint32_t)w * (uint32_t)h;
  if (npix == 0) return hipSuccess;
  hipLaunchKernelGGL(k_accumulate, dim3((npix + 255) / 256), dim3(256), 0, st, off, cnt, w, h, wt, val, spp, exposure,
                     rgb, u8);
  return hipGetLastError();
}

}  // namespace pm
