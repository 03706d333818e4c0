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

namespace pm {

// One wave (one workgroup) per 8x8 tile of the crop; lane = pixel.
template <bool FP64T, bool CONE>
__global__ void __launch_bounds__(64) k_render_lane(RenderLaunch r) {
  extern __shared__ uint32_t heap_lds[];  // [2][64 slots][64 lanes]
  const int lane = lane_id();
  const int tiles_x = (r.w + 7) / 8;
  const int tile = (int)blockIdx.x;
  const int tx = tile % tiles_x, ty = tile / tiles_x;
  const int lx = tx * 8 + (lane & 7), ly = ty * 8 + (lane >> 3);
  if (lx >= r.w || ly >= r.h) return;
  const uint32_t px = (uint32_t)(r.x0 + lx), py = (uint32_t)(r.y0 + ly);
  LaneGather<CONE> g;
  g.gs = r.gs;
  g.gv = r.gv;
  g.hp.H = heap_lds;
  g.hp.Lo = heap_lds + 64 * 64;
  g.hp.lane = lane;
  g.hp.k = 0;
  g.hp.size = 0;
  g.hg_g = r.sc.med.g;
  g.nsurf = 0u;
  g.nvol = 0u;
  g.prev_r = -1.0f;
  NSum3<float> acc;
  nsum3_init(acc);
  for (int s = 0; s < r.rp.spp; ++s) {
    V3f v = FP64T ? render_sample_t<float, double>(r.sc, r.rp, px, py, (uint32_t)s, g)
                  : render_sample_t<float, float>(r.sc, r.rp, px, py, (uint32_t)s, g);
    nsum3_add(acc, v);
  }
  V3f v = nsum3_get(acc) * (1.0f / (float)r.rp.spp);
  const size_t o = (size_t)ly * (size_t)r.w + (size_t)lx;
  if (r.rgb) {
    r.rgb[3 * o] = v.x;
    r.rgb[3 * o + 1] = v.y;
    r.rgb[3 * o + 2] = v.z;
  }
  if (r.u8) {
    uchar4 q;
    q.x = tonemap_u8(v.x, r.exposure);
    q.y = tonemap_u8(v.y, r.exposure);
    q.z = tonemap_u8(v.z, r.exposure);
    q.w = 0;
    r.u8[o] = q;
  }
  if (r.counters) {
    atomicAdd(&r.counters[0], (unsigned long long)g.nsurf);
    atomicAdd(&r.counters[1], (unsigned long long)g.nvol);
  }
}

hipError_t launch_render_lane(const RenderLaunch& r, hipStream_t st) {
  if (r.w <= 0 || r.h <= 0) return hipSuccess;
  unsigned tiles = (unsigned)(((r.w + 7) / 8) * ((r.h + 7) / 8));
  size_t lds = 2 * 64 * 64 * sizeof(uint32_t);
  const bool fp64 = r.rp.fp64_transmittance != 0, cone = r.estimator == 1;
  if (fp64 && cone) hipLaunchKernelGGL((k_render_lane<true, true>), dim3(tiles), dim3(64), lds, st, r);
  else if (fp64) hipLaunchKernelGGL((k_render_lane<true, false>), dim3(tiles), dim3(64), lds, st, r);
  else if (cone) hipLaunchKernelGGL((k_render_lane<false, true>), dim3(tiles), dim3(64), lds, st, r);
  else hipLaunchKernelGGL((k_render_lane<false, false>), dim3(tiles), dim3(64), lds, st, r);
  return hipGetLastError();
}

}  // namespace pm
