// pm_gather_lane.hip -- render with one lane per pixel and a per-lane kNN.
//
// Same inputs, outputs and neighbour sets as pm_gather.hip (the reference
// counterpart is photon_mapping_kernel, photonMappingKernel.cu:1409-1462, with
// computePixelColor :926-1017 and the gathers :314-389, :831-870), organised
// for throughput instead of per-query latency:
//  * a wave renders an 8x8 pixel tile, one pixel per lane, so the 64 queries
//    in flight are spatially coherent (neighbouring pixels, same march step);
//  * each lane keeps its k best (FP32 bits of d^2, photon index) keys in a
//    bounded max-heap in LDS (k <= 64 slots x 8 B per lane, slot-major so the
//    lanes of a wave touch consecutive banks);
//  * traversal is stackless: near-first depth-first search over the
//    heap-ordered tree using the implicit parent (h-1)/2 and each internal
//    node's point range, so no per-lane stack memory is needed;
//  * the volume query along the march is seeded with the bound
//    (r_prev + |q - q_prev|)^2 from the same lane's previous query; when fewer
//    than k keys fall under it the query is repeated with the full radius cap,
//    so the result is always the exact kNN set.
#include "pm_lane_knn.h"
#include "pm_lane_knn2.h"
#include "pm_lane_knn3.h"
#include "pm_lane_knn4.h"
#include "pm_lane_knn5.h"

namespace pm {

// Radiance of pixel (px, py): the mean of its spp samples, compensated sum.
// heap: this wave's LDS heap, `slots` keys per lane (k <= slots).
template <bool FP64T, bool CONE>
__device__ __forceinline__ V3f lane_pixel(const Scene& sc, const RenderParams& rp, const GatherMap& gs,
                                          const GatherMap& gv, uint32_t* heap, int slots, int lane, uint32_t px,
                                          uint32_t py, uint32_t& nsurf, uint32_t& nvol) {
  LaneGather<CONE> g;
  g.gs = gs;
  g.gv = gv;
  g.hp.H = heap;
  g.hp.Lo = heap + slots * 64;
  g.hp.lane = lane;
  g.hp.k = 0;
  g.hp.size = 0;
  g.hg_g = sc.med.g;
  g.nsurf = 0u;
  g.nvol = 0u;
  g.prev_r = -1.0f;
  NSum3<float> acc;
  nsum3_init(acc);
  for (int s = 0; s < rp.spp; ++s) {
    V3f v = FP64T ? render_sample_t<float, double>(sc, rp, px, py, (uint32_t)s, g)
                  : render_sample_t<float, float>(sc, rp, px, py, (uint32_t)s, g);
    nsum3_add(acc, v);
  }
  nsurf = g.nsurf;
  nvol = g.nvol;
  return nsum3_get(acc) * (1.0f / (float)rp.spp);
}

// The same with the 4-byte-heap gather (pm_lane_knn2.h); heap: `64` slots of
// 4 bytes per lane.
template <bool FP64T, bool CONE>
__device__ __forceinline__ V3f lane_pixel2(const Scene& sc, const RenderParams& rp, const GatherMap& gs,
                                           const GatherMap& gv, uint32_t* heap, int lane, uint32_t px, uint32_t py,
                                           uint32_t& nsurf, uint32_t& nvol) {
  LaneGather2<CONE> g;
  g.gs = gs;
  g.gv = gv;
  g.hp.H = heap;
  g.hp.lane = lane;
  g.hp.k = 0;
  g.hp.size = 0;
  g.hg_g = sc.med.g;
  g.nsurf = 0u;
  g.nvol = 0u;
  g.prev_r = -1.0f;
  NSum3<float> acc;
  nsum3_init(acc);
  for (int s = 0; s < rp.spp; ++s) {
    V3f v = FP64T ? render_sample_t<float, double>(sc, rp, px, py, (uint32_t)s, g)
                  : render_sample_t<float, float>(sc, rp, px, py, (uint32_t)s, g);
    nsum3_add(acc, v);
  }
  nsurf = g.nsurf;
  nvol = g.nvol;
  return nsum3_get(acc) * (1.0f / (float)rp.spp);
}

// The same with the 2-byte-heap gather (pm_lane_knn3.h).
template <bool FP64T, bool CONE, int NB = 8>
__device__ __forceinline__ V3f lane_pixel3(const Scene& sc, const RenderParams& rp, const GatherMap& gs,
                                           const GatherMap& gv, uint16_t* heap, int lane, uint32_t px, uint32_t py,
                                           uint32_t& nsurf, uint32_t& nvol) {
  LaneGather3<CONE, NB> g;
  g.gs = gs;
  g.gv = gv;
  g.hp.H = heap;
  g.hp.lane = lane;
  g.hp.k = 0;
  g.hp.size = 0;
  g.hg_g = sc.med.g;
  g.nsurf = 0u;
  g.nvol = 0u;
  g.prev_r = -1.0f;
  NSum3<float> acc;
  nsum3_init(acc);
  for (int s = 0; s < rp.spp; ++s) {
    V3f v = FP64T ? render_sample_t<float, double>(sc, rp, px, py, (uint32_t)s, g)
                  : render_sample_t<float, float>(sc, rp, px, py, (uint32_t)s, g);
    nsum3_add(acc, v);
  }
  nsurf = g.nsurf;
  nvol = g.nvol;
  return nsum3_get(acc) * (1.0f / (float)rp.spp);
}

// The same with the deferred-volume gathers (pm_lane_knn4.h, pm_lane_knn5.h).
template <bool FP64T, bool CONE, class G = LaneGather4<CONE, 4>>
__device__ __forceinline__ V3f lane_pixel4(const Scene& sc, const RenderParams& rp, const GatherMap& gs,
                                           const GatherMap& gv, uint16_t* heap, int lane, uint32_t px, uint32_t py,
                                           uint32_t& nsurf, uint32_t& nvol) {
  G g;
  g.gs = gs;
  g.gv = gv;
  g.hp.H = heap;
  g.hp.lane = lane;
  g.hp.k = 0;
  g.hp.size = 0;
  g.hg_g = sc.med.g;
  g.nsurf = 0u;
  g.nvol = 0u;
  g.prev_r = -1.0f;
  g.pend = false;
  g.pend_full = false;
  NSum3<float> acc;
  nsum3_init(acc);
  for (int s = 0; s < rp.spp; ++s) {
    V3f v = FP64T ? render_sample_t<float, double>(sc, rp, px, py, (uint32_t)s, g)
                  : render_sample_t<float, float>(sc, rp, px, py, (uint32_t)s, g);
    nsum3_add(acc, v);
  }
  nsurf = g.nsurf;
  nvol = g.nvol;
  return nsum3_get(acc) * (1.0f / (float)rp.spp);
}

__device__ __forceinline__ void store_pixel(float* rgb, uchar4* u8, size_t o, V3f v, float exposure) {
  if (rgb) {
    rgb[3 * o] = v.x;
    rgb[3 * o + 1] = v.y;
    rgb[3 * o + 2] = v.z;
  }
  if (u8) {
    uchar4 q;
    q.x = tonemap_u8(v.x, exposure);
    q.y = tonemap_u8(v.y, exposure);
    q.z = tonemap_u8(v.z, exposure);
    q.w = 0;
    u8[o] = q;
  }
}

// Workgroups are dispatched to the 8 XCDs of the chip round-robin, and each
// XCD has its own L2.  With tile_regions = rx (1, 2, 4 or 8) the crop's tiles
// are cut into an rx x (8 / rx) grid of regions and workgroup b renders the
// (b / 8)-th tile, in raster order, of region b % 8: each XCD then works on
// one compact part of the frame, and its L2 holds the photons of that part's
// view frustum.  tile_regions = 0 renders tile b in raster order.
__host__ __device__ __forceinline__ void lane_regions(const RenderLaunch& r, int& tiles_x, int& tiles_y, int& rx,
                                                      int& ry, int& rw, int& rh) {
  tiles_x = (r.w + 7) / 8;
  tiles_y = (r.h + 7) / 8;
  rx = r.tile_regions;
  ry = 8 / (rx > 0 ? rx : 1);
  rw = (tiles_x + rx - 1) / (rx > 0 ? rx : 1);
  rh = (tiles_y + ry - 1) / ry;
}

unsigned lane_grid(const RenderLaunch& r) {
  int tiles_x, tiles_y, rx, ry, rw, rh;
  lane_regions(r, tiles_x, tiles_y, rx, ry, rw, rh);
  if (rx <= 0) return (unsigned)(tiles_x * tiles_y);
  return 8u * (unsigned)(rw * rh);
}

// Pixel (lx, ly) of the crop for this lane; false when it lies outside.
__device__ __forceinline__ bool lane_tile(const RenderLaunch& r, int lane, int& lx, int& ly) {
  int tiles_x, tiles_y, rx, ry, rw, rh;
  lane_regions(r, tiles_x, tiles_y, rx, ry, rw, rh);
  const int b = (int)blockIdx.x;
  int tx, ty;
  if (rx <= 0) {
    tx = b % tiles_x;
    ty = b / tiles_x;
  } else {
    const int reg = b & 7, slot = b >> 3;
    const int gx = reg % rx, gy = reg / rx;
    // region bounds: the first (tiles mod rx) regions of a row are one wider
    const int x0 = gx * (tiles_x / rx) + min(gx, tiles_x % rx), w = tiles_x / rx + (gx < tiles_x % rx ? 1 : 0);
    const int y0 = gy * (tiles_y / ry) + min(gy, tiles_y % ry), h = tiles_y / ry + (gy < tiles_y % ry ? 1 : 0);
    if (w <= 0 || slot >= w * h) return false;
    tx = x0 + slot % w;
    ty = y0 + slot / w;
  }
  lx = tx * 8 + (lane & 7);
  ly = ty * 8 + (lane >> 3);
  return lx < r.w && ly < r.h;
}

// One wave (one workgroup) per 8x8 tile of the crop; lane = pixel.
template <bool FP64T, bool CONE>
__global__ void __launch_bounds__(64) k_render_lane(RenderLaunch r) {
  extern __shared__ uint32_t heap_lds[];  // [2][64 slots][64 lanes]
  const int lane = lane_id();
  int lx, ly;
  if (!lane_tile(r, lane, lx, ly)) return;
  const uint32_t px = (uint32_t)(r.x0 + lx), py = (uint32_t)(r.y0 + ly);
  uint32_t ns, nv;
  V3f v = lane_pixel<FP64T, CONE>(r.sc, r.rp, r.gs, r.gv, heap_lds, 64, lane, px, py, ns, nv);
  store_pixel(r.rgb, r.u8, (size_t)ly * (size_t)r.w + (size_t)lx, v, r.exposure);
  if (r.counters) {
    atomicAdd(&r.counters[0], (unsigned long long)ns);
    atomicAdd(&r.counters[1], (unsigned long long)nv);
  }
}

// Gather mode 2: k_render_lane with the 4-byte heap (16 KB of LDS per wave).
template <bool FP64T, bool CONE>
__global__ void __launch_bounds__(64) k_render_lane2(RenderLaunch r) {
  extern __shared__ uint32_t heap_lds[];  // [64 slots][64 lanes]
  const int lane = lane_id();
  int lx, ly;
  if (!lane_tile(r, lane, lx, ly)) return;
  const uint32_t px = (uint32_t)(r.x0 + lx), py = (uint32_t)(r.y0 + ly);
  uint32_t ns, nv;
  V3f v = lane_pixel2<FP64T, CONE>(r.sc, r.rp, r.gs, r.gv, heap_lds, lane, px, py, ns, nv);
  store_pixel(r.rgb, r.u8, (size_t)ly * (size_t)r.w + (size_t)lx, v, r.exposure);
  if (r.counters) {
    atomicAdd(&r.counters[0], (unsigned long long)ns);
    atomicAdd(&r.counters[1], (unsigned long long)nv);
  }
}

// Gather mode 4: the 2-byte heap (8 KB of LDS per wave), leaf points loaded 4
// at a time, with the register budget of 5 waves per SIMD.
template <bool FP64T, bool CONE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) k_render_lane4(RenderLaunch r) {
  extern __shared__ uint16_t heap16_lds[];  // [64 slots][64 lanes]
  const int lane = lane_id();
  int lx, ly;
  if (!lane_tile(r, lane, lx, ly)) return;
  const uint32_t px = (uint32_t)(r.x0 + lx), py = (uint32_t)(r.y0 + ly);
  uint32_t ns, nv;
  V3f v = lane_pixel3<FP64T, CONE, 4>(r.sc, r.rp, r.gs, r.gv, heap16_lds, lane, px, py, ns, nv);
  store_pixel(r.rgb, r.u8, (size_t)ly * (size_t)r.w + (size_t)lx, v, r.exposure);
  if (r.counters) {
    atomicAdd(&r.counters[0], (unsigned long long)ns);
    atomicAdd(&r.counters[1], (unsigned long long)nv);
  }
}

hipError_t launch_render_lane4(const RenderLaunch& r, hipStream_t st) {
  if (r.w <= 0 || r.h <= 0) return hipSuccess;
  const unsigned tiles = lane_grid(r);
  size_t lds = 64 * 64 * sizeof(uint16_t);
  const bool fp64 = r.rp.fp64_transmittance != 0, cone = r.estimator == 1;
  if (fp64 && cone) hipLaunchKernelGGL((k_render_lane4<true, true>), dim3(tiles), dim3(64), lds, st, r);
  else if (fp64) hipLaunchKernelGGL((k_render_lane4<true, false>), dim3(tiles), dim3(64), lds, st, r);
  else if (cone) hipLaunchKernelGGL((k_render_lane4<false, true>), dim3(tiles), dim3(64), lds, st, r);
  else hipLaunchKernelGGL((k_render_lane4<false, false>), dim3(tiles), dim3(64), lds, st, r);
  return hipGetLastError();
}

// LDS of a mode-7 wave: the 8 KB of 2-byte heap keys / bucket counters, then
// with TOP the top-of-tree tiles of the surface and volume trees.
constexpr size_t kLane7HeapBytes = 64 * 64 * sizeof(uint16_t);

template <bool TOP>
__device__ __forceinline__ void stage_lane7_tiles(GatherMap& gs, GatherMap& gv, uint16_t* heap, int levels,
                                                  int lane) {
  if constexpr (TOP) {
    unsigned char* t = (unsigned char*)heap + kLane7HeapBytes;
    stage_top_tile(gs, t, levels, lane);
    stage_top_tile(gv, t + top_tile_bytes(levels), levels, lane);
  }
}

// Gather mode 7: the 2-byte-key gather of mode 4 with each march step's
// second pass merged into the next step's first-pass traversal (the deferred
// volume query of pm_lane_knn4.h) and the first passes done by counting in a
// bucket window (pm_lane_knn5.h), 5 waves per SIMD.  TOP: the traversals read
// the top r.top_levels levels of both trees from LDS.
template <bool FP64T, bool CONE, bool TOP>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) k_render_lane7(RenderLaunch r) {
  extern __shared__ uint16_t heap16_lds[];  // [64 slots][64 lanes], then the top tiles
  const int lane = lane_id();
  GatherMap gs = r.gs, gv = r.gv;
  stage_lane7_tiles<TOP>(gs, gv, heap16_lds, r.top_levels, lane);  // by every lane, before any leaves
  int lx, ly;
  if (!lane_tile(r, lane, lx, ly)) return;
  const uint32_t px = (uint32_t)(r.x0 + lx), py = (uint32_t)(r.y0 + ly);
  uint32_t ns, nv;
  V3f v = lane_pixel4<FP64T, CONE, LaneGather5<CONE, false, TOP>>(r.sc, r.rp, gs, gv, heap16_lds, lane, px, py, ns,
                                                                   nv);
  store_pixel(r.rgb, r.u8, (size_t)ly * (size_t)r.w + (size_t)lx, v, r.exposure);
  if (r.counters) {
    atomicAdd(&r.counters[0], (unsigned long long)ns);
    atomicAdd(&r.counters[1], (unsigned long long)nv);
  }
}

// Mode 7 with the samples of a pixel in adjacent lanes.  For spp = 2, 4, 8 or
// 16 a wave renders a tile of 64 / spp pixels (8x4, 4x4, 4x2 or 2x2), and lane
// p * spp + s renders sample s of pixel p on its own; the pixel is the
// compensated sum of its samples in order s = 0, 1, ..., as in the one-lane
// sample loop.  The lanes of a pixel march nearly the same ray, so their
// traversals and leaf batches line up.
__device__ __forceinline__ int sample_tile_w(int spp) {
  const int P = 64 / spp;
  return P == 32 ? 8 : (P == 16 ? 4 : (P == 8 ? 4 : 2));
}

template <bool FP64T, bool CONE, bool TOP>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) k_render_lane7s(RenderLaunch r) {
  extern __shared__ uint16_t heap16_lds[];  // [64 slots][64 lanes], then the top tiles
  const int lane = lane_id();
  LaneGather5<CONE, false, TOP> g;
  g.gs = r.gs;
  g.gv = r.gv;
  stage_lane7_tiles<TOP>(g.gs, g.gv, heap16_lds, r.top_levels, lane);  // by every lane, before any leaves
  const int spp = r.rp.spp, lg = __builtin_ctz((unsigned)spp);
  const int pw = sample_tile_w(spp), ph = (64 / spp) / pw;
  const int p = lane >> lg, s = lane & (spp - 1);
  const int tiles_x = (r.w + pw - 1) / pw;
  const int tx = (int)blockIdx.x % tiles_x, ty = (int)blockIdx.x / tiles_x;
  const int lx = tx * pw + p % pw, ly = ty * ph + p / pw;
  if (lx >= r.w || ly >= r.h) return;  // all lanes of a pixel leave together
  const uint32_t px = (uint32_t)(r.x0 + lx), py = (uint32_t)(r.y0 + ly);
  g.hp.H = heap16_lds;
  g.hp.lane = lane;
  g.hp.k = 0;
  g.hp.size = 0;
  g.hg_g = r.sc.med.g;
  g.nsurf = 0u;
  g.nvol = 0u;
  g.prev_r = -1.0f;
  g.pend = false;
  g.pend_full = false;
  const V3f v = FP64T ? render_sample_t<float, double>(r.sc, r.rp, px, py, (uint32_t)s, g)
                      : render_sample_t<float, float>(r.sc, r.rp, px, py, (uint32_t)s, g);
  NSum3<float> acc;
  nsum3_init(acc);
  const int base = p << lg;
  for (int t = 0; t < spp; ++t)
    nsum3_add(acc, v3(lane_read_f(v.x, base + t), lane_read_f(v.y, base + t), lane_read_f(v.z, base + t)));
  if (s == 0) store_pixel(r.rgb, r.u8, (size_t)ly * (size_t)r.w + (size_t)lx, nsum3_get(acc) * (1.0f / (float)spp),
                          r.exposure);
  if (r.counters) {
    atomicAdd(&r.counters[0], (unsigned long long)g.nsurf);
    atomicAdd(&r.counters[1], (unsigned long long)g.nvol);
  }
}

template <bool TOP>
static hipError_t launch_render_lane7_t(const RenderLaunch& r, hipStream_t st) {
  const size_t lds = kLane7HeapBytes + (TOP ? 2 * top_tile_bytes(r.top_levels) : 0);
  const bool fp64 = r.rp.fp64_transmittance != 0, cone = r.estimator == 1;
  const int spp = r.rp.spp;
  if (r.sample_lanes && (spp == 2 || spp == 4 || spp == 8 || spp == 16)) {
    const int pw = spp == 4 || spp == 8 ? 4 : (spp == 2 ? 8 : 2), ph = (64 / spp) / pw;
    const unsigned tiles = (unsigned)(((r.w + pw - 1) / pw) * ((r.h + ph - 1) / ph));
    if (fp64 && cone) hipLaunchKernelGGL((k_render_lane7s<true, true, TOP>), dim3(tiles), dim3(64), lds, st, r);
    else if (fp64) hipLaunchKernelGGL((k_render_lane7s<true, false, TOP>), dim3(tiles), dim3(64), lds, st, r);
    else if (cone) hipLaunchKernelGGL((k_render_lane7s<false, true, TOP>), dim3(tiles), dim3(64), lds, st, r);
    else hipLaunchKernelGGL((k_render_lane7s<false, false, TOP>), dim3(tiles), dim3(64), lds, st, r);
    return hipGetLastError();
  }
  const unsigned tiles = lane_grid(r);
  if (fp64 && cone) hipLaunchKernelGGL((k_render_lane7<true, true, TOP>), dim3(tiles), dim3(64), lds, st, r);
  else if (fp64) hipLaunchKernelGGL((k_render_lane7<true, false, TOP>), dim3(tiles), dim3(64), lds, st, r);
  else if (cone) hipLaunchKernelGGL((k_render_lane7<false, true, TOP>), dim3(tiles), dim3(64), lds, st, r);
  else hipLaunchKernelGGL((k_render_lane7<false, false, TOP>), dim3(tiles), dim3(64), lds, st, r);
  return hipGetLastError();
}

hipError_t launch_render_lane7(const RenderLaunch& r, hipStream_t st) {
  if (r.w <= 0 || r.h <= 0) return hipSuccess;
  if (r.top_levels < 0 || r.top_levels > kMaxTopLevels) return hipErrorInvalidValue;
  return r.top_levels > 0 ? launch_render_lane7_t<true>(r, st) : launch_render_lane7_t<false>(r, st);
}

hipError_t launch_render_lane2(const RenderLaunch& r, hipStream_t st) {
  if (r.w <= 0 || r.h <= 0) return hipSuccess;
  const unsigned tiles = lane_grid(r);
  size_t lds = 64 * 64 * sizeof(uint32_t);
  const bool fp64 = r.rp.fp64_transmittance != 0, cone = r.estimator == 1;
  if (fp64 && cone) hipLaunchKernelGGL((k_render_lane2<true, true>), dim3(tiles), dim3(64), lds, st, r);
  else if (fp64) hipLaunchKernelGGL((k_render_lane2<true, false>), dim3(tiles), dim3(64), lds, st, r);
  else if (cone) hipLaunchKernelGGL((k_render_lane2<false, true>), dim3(tiles), dim3(64), lds, st, r);
  else hipLaunchKernelGGL((k_render_lane2<false, false>), dim3(tiles), dim3(64), lds, st, r);
  return hipGetLastError();
}

hipError_t launch_render_lane(const RenderLaunch& r, hipStream_t st) {
  if (r.w <= 0 || r.h <= 0) return hipSuccess;
  const unsigned tiles = lane_grid(r);
  size_t lds = 2 * 64 * 64 * sizeof(uint32_t);
  const bool fp64 = r.rp.fp64_transmittance != 0, cone = r.estimator == 1;
  if (fp64 && cone) hipLaunchKernelGGL((k_render_lane<true, true>), dim3(tiles), dim3(64), lds, st, r);
  else if (fp64) hipLaunchKernelGGL((k_render_lane<true, false>), dim3(tiles), dim3(64), lds, st, r);
  else if (cone) hipLaunchKernelGGL((k_render_lane<false, true>), dim3(tiles), dim3(64), lds, st, r);
  else hipLaunchKernelGGL((k_render_lane<false, false>), dim3(tiles), dim3(64), lds, st, r);
  return hipGetLastError();
}

// Batched render: one persistent grid over the 8x8 tiles of every scene of a
// batch.  Each wave takes the next (scene, tile) item from a device work
// queue (an atomic counter over an item list ordered by decreasing expected
// cost) until the list is exhausted, so scenes of very different sizes keep
// all waves busy.  Per pixel the arithmetic is that of the one-scene lane
// render with the same gather (G = 8: mode 1, 4: mode 2, 2: mode 4, 7: mode 7
// with 2-byte keys), so every scene's image is bit-identical to its own render
// in that mode.  heap_slots (32 or 64 keys per lane) and the key width size
// each wave's LDS heap; mode 7 counts in the same 8 KB (64 slots).
// TOP (gather mode 7 only): each wave stages the top r.top_levels levels of
// its current scene's two trees in its own LDS region after the heaps, and
// stages them again when its next work item belongs to another scene.
template <bool FP64T, bool CONE, int G, bool TOP>
__global__ void __launch_bounds__(256) k_render_batch(BatchRenderLaunch r) {
  constexpr int KB = G == 7 ? 2 : G;
  static_assert(!TOP || G == 7, "top-of-tree tiles are read by the mode-7 gather");
  extern __shared__ uint32_t lds[];
  const int lane = lane_id();
  const int wave = (int)uniform(threadIdx.x >> 6);  // wave-uniform: the LDS regions' addresses stay scalar
  const size_t heap_words = (size_t)r.heap_slots * 64u * KB / 4u;
  uint32_t* heap = lds + (size_t)wave * heap_words;
  unsigned char* tiles = nullptr;
  GatherMap tgs, tgv;
  uint32_t staged = 0xffffffffu;
  if constexpr (TOP)
    tiles = (unsigned char*)(lds + (size_t)(blockDim.x >> 6) * heap_words) +
            (size_t)wave * 2u * top_tile_bytes(r.top_levels);
  unsigned long long ns_tot = 0, nv_tot = 0;
  for (;;) {
    uint32_t it = 0;
    if (lane == 0) it = atomicAdd(r.head, 1u);
    it = read_lane(it, 0);
    if (it >= r.nitems) break;
    const uint2 item = r.items[it];
    // wave-uniform scene and tile: scalar loads of the scene descriptor
    const uint32_t scene = uniform(item.x);
    const BatchRenderScene& S = r.scenes[scene];
    if constexpr (TOP) {
      if (scene != staged) {  // every lane of the wave is here
        tgs = S.gs;
        tgv = S.gv;
        stage_top_tile(tgs, tiles, r.top_levels, lane);
        stage_top_tile(tgv, tiles + top_tile_bytes(r.top_levels), r.top_levels, lane);
        staged = scene;
      }
    }
    const uint32_t tile = uniform(item.y);
    const int tiles_x = (S.w + 7) / 8;
    const int tx = (int)tile % tiles_x, ty = (int)tile / tiles_x;
    const int lx = tx * 8 + (lane & 7), ly = ty * 8 + (lane >> 3);
    if (lx < S.w && ly < S.h) {
      uint32_t ns, nv;
      V3f v;
      if (KB == 8)
        v = lane_pixel<FP64T, CONE>(S.sc, S.rp, S.gs, S.gv, heap, r.heap_slots, lane, (uint32_t)lx, (uint32_t)ly, ns,
                                    nv);
      else if (KB == 4)
        v = lane_pixel2<FP64T, CONE>(S.sc, S.rp, S.gs, S.gv, heap, lane, (uint32_t)lx, (uint32_t)ly, ns, nv);
      else if (G == 7 && TOP)
        v = lane_pixel4<FP64T, CONE, LaneGather5<CONE, false, true>>(S.sc, S.rp, tgs, tgv, (uint16_t*)heap, lane,
                                                                     (uint32_t)lx, (uint32_t)ly, ns, nv);
      else if (G == 7)
        v = lane_pixel4<FP64T, CONE, LaneGather5<CONE>>(S.sc, S.rp, S.gs, S.gv, (uint16_t*)heap, lane, (uint32_t)lx,
                                                        (uint32_t)ly, ns, nv);
      else
        v = lane_pixel3<FP64T, CONE, 4>(S.sc, S.rp, S.gs, S.gv, (uint16_t*)heap, lane, (uint32_t)lx, (uint32_t)ly, ns,
                                        nv);
      store_pixel(r.rgb + 3 * S.img_off, nullptr, (size_t)ly * (size_t)S.w + (size_t)lx, v, S.exposure);
      ns_tot += ns;
      nv_tot += nv;
    }
  }
  if (r.counters) {
    atomicAdd(&r.counters[0], ns_tot);
    atomicAdd(&r.counters[1], nv_tot);
  }
}

size_t render_batch_lds_bytes(int block_threads, int heap_slots, int key_bytes, int top_levels) {
  return (size_t)(block_threads / 64) * ((size_t)heap_slots * 64u * (size_t)key_bytes + 2u * top_tile_bytes(top_levels));
}

// Gather kind of a launch: the heap key width, or 7 for mode 7.
static int batch_gather(const BatchRenderLaunch& r) { return r.counting ? 7 : r.key_bytes; }

template <int KB, bool TOP>
static hipError_t launch_render_batch_kb(const BatchRenderLaunch& r, int fp64, int cone, int block_threads, int grid,
                                         hipStream_t st) {
  const size_t lds = render_batch_lds_bytes(block_threads, r.heap_slots, KB == 7 ? 2 : KB, TOP ? r.top_levels : 0);
  if (fp64 && cone)
    hipLaunchKernelGGL((k_render_batch<true, true, KB, TOP>), dim3(grid), dim3(block_threads), lds, st, r);
  else if (fp64)
    hipLaunchKernelGGL((k_render_batch<true, false, KB, TOP>), dim3(grid), dim3(block_threads), lds, st, r);
  else if (cone)
    hipLaunchKernelGGL((k_render_batch<false, true, KB, TOP>), dim3(grid), dim3(block_threads), lds, st, r);
  else
    hipLaunchKernelGGL((k_render_batch<false, false, KB, TOP>), dim3(grid), dim3(block_threads), lds, st, r);
  return hipGetLastError();
}

hipError_t launch_render_batch(const BatchRenderLaunch& r, int fp64, int cone, int block_threads, int grid,
                               hipStream_t st) {
  if (r.nitems == 0) return hipSuccess;
  const int G = batch_gather(r);
  if (G == 7 && (r.key_bytes != 2 || r.heap_slots != 64)) return hipErrorInvalidValue;
  if (r.top_levels < 0 || r.top_levels > kMaxTopLevels || (r.top_levels > 0 && G != 7)) return hipErrorInvalidValue;
  if (G == 7 && r.top_levels > 0) return launch_render_batch_kb<7, true>(r, fp64, cone, block_threads, grid, st);
  if (G == 7) return launch_render_batch_kb<7, false>(r, fp64, cone, block_threads, grid, st);
  if (G == 8) return launch_render_batch_kb<8, false>(r, fp64, cone, block_threads, grid, st);
  if (G == 4) return launch_render_batch_kb<4, false>(r, fp64, cone, block_threads, grid, st);
  return launch_render_batch_kb<2, false>(r, fp64, cone, block_threads, grid, st);
}

template <int KB, bool TOP>
static int occupancy_kb(int fp64, int cone, int block_threads, size_t lds) {
  int nb = 0;
  hipError_t e;
  if (fp64 && cone)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render_batch<true, true, KB, TOP>, block_threads, lds);
  else if (fp64)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render_batch<true, false, KB, TOP>, block_threads, lds);
  else if (cone)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render_batch<false, true, KB, TOP>, block_threads, lds);
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render_batch<false, false, KB, TOP>, block_threads, lds);
  return e == hipSuccess ? nb : 0;
}

int render_batch_occupancy(int fp64, int cone, int block_threads, int heap_slots, int key_bytes, int counting,
                           int top_levels) {
  const size_t lds = render_batch_lds_bytes(block_threads, heap_slots, key_bytes, counting ? top_levels : 0);
  if (counting && top_levels > 0) return occupancy_kb<7, true>(fp64, cone, block_threads, lds);
  if (counting) return occupancy_kb<7, false>(fp64, cone, block_threads, lds);
  if (key_bytes == 8) return occupancy_kb<8, false>(fp64, cone, block_threads, lds);
  if (key_bytes == 4) return occupancy_kb<4, false>(fp64, cone, block_threads, lds);
  return occupancy_kb<2, false>(fp64, cone, block_threads, lds);
}

__global__ void k_tonemap(const float* __restrict__ rgb, size_t npix, float exposure, uchar4* __restrict__ out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  uchar4 q;
  q.x = tonemap_u8(rgb[3 * i], exposure);
  q.y = tonemap_u8(rgb[3 * i + 1], exposure);
  q.z = tonemap_u8(rgb[3 * i + 2], exposure);
  q.w = 0;
  out[i] = q;
}

hipError_t launch_tonemap(const float* rgb, size_t npix, float exposure, uchar4* out, hipStream_t st) {
  if (npix == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tonemap, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, st, rgb, npix, exposure, out);
  return hipGetLastError();
}

}  // namespace pm
