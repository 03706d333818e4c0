// pm_voxel.hip -- the voxel-grid compatibility mode on gfx950: the photon
// maps splatted into 32^3 fixed-point grids, their summed-volume tables, and
// the render with the window estimates of csrc/core/pm_voxel.h.
//
// Reference counterparts: storePhoton / storeVolumePhoton
// (photonMappingKernel.cu:1147-1183, a racy float += into the voxel), the
// window gathers :314-389, :442-600, :831-870 and the pixel kernel
// photon_mapping_kernel :1409-1462.  Here the splat is one lane per record
// with 64-bit integer atomics (exact and order-independent), the tables are
// exact integer prefix sums, and a window sum is 8 table reads.
#include "pm_kernels.h"
#include "pm_wave.h"

#include "../core/pm_voxel.h"

namespace pm {

// One lane per record: its power (FP32 or BF16 payload) in fixed point, added
// to the record's voxel.
__global__ void k_voxel_splat(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                              const void* __restrict__ payload, int bf16, uint32_t n,
                              unsigned long long* __restrict__ grid) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float p[3];
  if (bf16) {
    const uint16_t* q = (const uint16_t*)payload + 8ull * i;
    for (int k = 0; k < 3; ++k) p[k] = bf16_to_f32(q[k]);
  } else {
    const float* q = (const float*)payload + 8ull * i;
    for (int k = 0; k < 3; ++k) p[k] = q[k];
  }
  const size_t cell = vox_cell(vox_of(x[i]), vox_of(y[i]), vox_of(z[i]));
  for (int k = 0; k < 3; ++k) atomicAdd(&grid[3 * cell + k], (unsigned long long)vox_fix(p[k]));
}

// Summed-volume tables of the two grids (blockIdx.y = map), first pass: one
// workgroup per table plane z, which it builds in LDS from grid plane z - 1
// (plane 0 and the first row and column of every plane are zero) with
// prefix sums along x, then y, and writes to the table.
__global__ void __launch_bounds__(256) k_voxel_sat_xy(const unsigned long long* __restrict__ grid_s,
                                                      const unsigned long long* __restrict__ grid_v,
                                                      unsigned long long* __restrict__ sat_s,
                                                      unsigned long long* __restrict__ sat_v) {
  __shared__ unsigned long long P[kSat * kSat * 3];
  const unsigned long long* grid = blockIdx.y ? grid_v : grid_s;
  unsigned long long* S = blockIdx.y ? sat_v : sat_s;
  const int z = (int)blockIdx.x, t = (int)threadIdx.x;
  for (int e = t; e < kSat * kSat * 3; e += 256) {
    const int ch = e % 3, xx = (e / 3) % kSat, yy = e / (3 * kSat);
    P[e] = (z && xx && yy) ? grid[3 * vox_cell(xx - 1, yy - 1, z - 1) + ch] : 0ull;
  }
  __syncthreads();
  if (t < kSat * 3) {  // prefix along x: one thread per (y, channel)
    const int ch = t % 3, yy = t / 3;
    unsigned long long run = 0ull;
    for (int xx = 0; xx < kSat; ++xx) {
      const int e = 3 * (yy * kSat + xx) + ch;
      run += P[e];
      P[e] = run;
    }
  }
  __syncthreads();
  if (t < kSat * 3) {  // prefix along y: one thread per (x, channel)
    const int ch = t % 3, xx = t / 3;
    unsigned long long run = 0ull;
    for (int yy = 0; yy < kSat; ++yy) {
      const int e = 3 * (yy * kSat + xx) + ch;
      run += P[e];
      P[e] = run;
    }
  }
  __syncthreads();
  for (int e = t; e < kSat * kSat * 3; e += 256) S[(size_t)z * kSat * kSat * 3 + e] = P[e];
}

// Second pass: prefix sums along z, one lane per (x, y, channel) line of
// either table (lanes [0, L) the surface table, [L, 2L) the volume table).
__global__ void __launch_bounds__(256) k_voxel_sat_z(unsigned long long* __restrict__ sat_s,
                                                     unsigned long long* __restrict__ sat_v) {
  constexpr int L = kSat * kSat * 3;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= 2 * L) return;
  unsigned long long* S = i < L ? sat_s : sat_v;
  const int e = i < L ? i : i - L;
  unsigned long long run = 0ull;
  for (int z = 0; z < kSat; ++z) {
    run += S[(size_t)z * L + e];
    S[(size_t)z * L + e] = run;
  }
}

// One lane per pixel of the crop: the eye samples with the voxel-grid
// estimates, mean of the spp samples by compensated sum.
template <bool FP64T>
__global__ void __launch_bounds__(64) k_render_voxel(RenderLaunch r, VoxelGrids G) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t ns = 0u, nv = 0u;
  if (i < (uint32_t)(r.w * r.h)) {
    const int lx = (int)(i % (uint32_t)r.w), ly = (int)(i / (uint32_t)r.w);
    const uint32_t px = (uint32_t)(r.x0 + lx), py = (uint32_t)(r.y0 + ly);
    VoxelGather<float> g;
    g.G = G;
    g.nsurf = 0u;
    g.nvol = 0u;
    NSum3<float> acc;
    nsum3_init(acc);
    for (int s = 0; s < r.rp.spp; ++s) {
      V3f v = FP64T ? render_sample_t<float, double>(r.sc, r.rp, px, py, (uint32_t)s, g)
                    : render_sample_t<float, float>(r.sc, r.rp, px, py, (uint32_t)s, g);
      nsum3_add(acc, v);
    }
    const V3f v = nsum3_get(acc) * (1.0f / (float)r.rp.spp);
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
    ns = g.nsurf;
    nv = g.nvol;
  }
  if (r.counters && (ns | nv)) {
    atomicAdd(&r.counters[0], (unsigned long long)ns);
    atomicAdd(&r.counters[1], (unsigned long long)nv);
  }
}

hipError_t launch_voxel_grids(const DevMap& ms, const DevMap& mv, unsigned long long* grid_s,
                              unsigned long long* grid_v, unsigned long long* sat_s, unsigned long long* sat_v,
                              hipStream_t st) {
  const size_t words = 3ull * kVox * kVox * kVox;
  hipError_t e = launch_fill_u32((uint32_t*)grid_s, 0u, 2 * words, st);
  if (e == hipSuccess) e = launch_fill_u32((uint32_t*)grid_v, 0u, 2 * words, st);
  if (e != hipSuccess) return e;
  if (ms.n)
    hipLaunchKernelGGL(k_voxel_splat, dim3((ms.n + 255) / 256), dim3(256), 0, st, ms.x, ms.y, ms.z, ms.payload,
                       ms.bf16, ms.n, grid_s);
  if (mv.n)
    hipLaunchKernelGGL(k_voxel_splat, dim3((mv.n + 255) / 256), dim3(256), 0, st, mv.x, mv.y, mv.z, mv.payload,
                       mv.bf16, mv.n, grid_v);
  hipLaunchKernelGGL(k_voxel_sat_xy, dim3(kSat, 2), dim3(256), 0, st, grid_s, grid_v, sat_s, sat_v);
  constexpr int lines = 2 * kSat * kSat * 3;
  hipLaunchKernelGGL(k_voxel_sat_z, dim3((lines + 255) / 256), dim3(256), 0, st, sat_s, sat_v);
  return hipGetLastError();
}

hipError_t launch_render_voxel(const RenderLaunch& r, const VoxelGrids& G, hipStream_t st) {
  const uint32_t n = (uint32_t)(r.w * r.h);
  const dim3 grid((n + 63) / 64);
  if (r.rp.fp64_transmittance) hipLaunchKernelGGL(k_render_voxel<true>, grid, dim3(64), 0, st, r, G);
  else hipLaunchKernelGGL(k_render_voxel<false>, grid, dim3(64), 0, st, r, G);
  return hipGetLastError();
}

}  // namespace pm
