// pm_scan_sort.hip -- exclusive scan and stable LSD radix sort (32-bit keys,
// 32-bit values) for gfx950.  Used by the photon-record compaction and the
// kd-tree build.  The reference has no counterpart (it never sorts or compacts:
// photons are splatted in place, photonMappingKernel.cu:1059-1183).
//
// Both are deterministic: the scan is exact integer arithmetic and the sort
// is stable (equal keys keep their input order), so sorting (key, index)
// pairs that arrive in index order yields the (key, index) order.
#include "pm_kernels.h"
#include "pm_wave.h"

namespace pm {

static constexpr int SCAN_THREADS = 256;
static constexpr int SCAN_ITEMS = 4;
static constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

// Block-wide exclusive scan of one value per thread (256 threads = 4 waves).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds_waves, uint32_t* total) {
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = lane_read(x, lane >= d ? lane - d : lane);
    if (lane >= d) x += y;
  }
  if (lane == 63) lds_waves[wave] = x;
  __syncthreads();
  uint32_t wave_off = 0, tot = 0;
  for (int w = 0; w < SCAN_THREADS / 64; ++w) {
    uint32_t s = lds_waves[w];
    if (w < wave) wave_off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return wave_off + x - v;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce(const uint32_t* __restrict__ in, size_t n,
                                                               uint32_t* __restrict__ partial) {
  __shared__ uint32_t lw[SCAN_THREADS / 64];
  size_t base = (size_t)blockIdx.x * SCAN_TILE;
  uint32_t s = 0;
  for (int r = 0; r < SCAN_ITEMS; ++r) {
    size_t i = base + (size_t)r * SCAN_THREADS + threadIdx.x;
    if (i < n) s += in[i];
  }
  uint32_t tot;
  block_exclusive_scan(s, lw, &tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Single-block exclusive scan of `n` partials in place; writes the grand total.
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_partials(uint32_t* __restrict__ partial, size_t n,
                                                                 uint32_t* __restrict__ total_out) {
  __shared__ uint32_t lw[SCAN_THREADS / 64];
  uint32_t carry = 0;
  for (size_t base = 0; base < n; base += SCAN_THREADS) {
    size_t i = base + threadIdx.x;
    uint32_t v = i < n ? partial[i] : 0u;
    uint32_t tot;
    uint32_t ex = block_exclusive_scan(v, lw, &tot);
    if (i < n) partial[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(const uint32_t* __restrict__ in, size_t n,
                                                              const uint32_t* __restrict__ partial,
                                                              uint32_t* __restrict__ out) {
  __shared__ uint32_t lw[SCAN_THREADS / 64];
  size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t s = 0;
  for (int r = 0; r < SCAN_ITEMS; ++r) {
    size_t i = base + r;
    v[r] = i < n ? in[i] : 0u;
    s += v[r];
  }
  uint32_t tot;
  uint32_t ex = block_exclusive_scan(s, lw, &tot) + partial[blockIdx.x];
  for (int r = 0; r < SCAN_ITEMS; ++r) {
    size_t i = base + r;
    if (i < n) out[i] = ex;
    ex += v[r];
  }
}

// One workgroup for a short array: tiles of SCAN_TILE in order with a
// running carry (the kd-build's per-level block counts, a few thousand).
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_small(const uint32_t* __restrict__ in, size_t n,
                                                              uint32_t* __restrict__ out,
                                                              uint32_t* __restrict__ total_out) {
  __shared__ uint32_t lw[SCAN_THREADS / 64];
  uint32_t carry = 0;
  for (size_t tile = 0; tile < n; tile += SCAN_TILE) {
    const size_t base = tile + (size_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
    for (int r = 0; r < SCAN_ITEMS; ++r) {
      v[r] = base + r < n ? in[base + r] : 0u;
      s += v[r];
    }
    uint32_t tot;
    uint32_t ex = carry + block_exclusive_scan(s, lw, &tot);
    for (int r = 0; r < SCAN_ITEMS; ++r) {
      if (base + r < n) out[base + r] = ex;
      ex += v[r];
    }
    carry += tot;
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}

size_t scan_temp_words(size_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, size_t n, uint32_t* temp, uint32_t* total_dev,
                          hipStream_t st) {
  if (n == 0) {
    if (total_dev) return hipMemsetAsync(total_dev, 0, 4, st);
    return hipSuccess;
  }
  size_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb <= 16) {  // one launch instead of three
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(SCAN_THREADS), 0, st, in, n, out, total_dev);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, st, in, n, temp);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(SCAN_THREADS), 0, st, temp, nb, total_dev);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, st, in, n, temp, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------- radix sort
static constexpr int RS_THREADS = 256;
static constexpr int RS_ROUNDS = 4;
static constexpr int RS_TILE = RS_THREADS * RS_ROUNDS;

__global__ void __launch_bounds__(RS_THREADS) k_radix_hist(const uint32_t* __restrict__ keys, size_t n, int shift,
                                                            uint32_t* __restrict__ hist, uint32_t nblocks) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  size_t base = (size_t)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ROUNDS; ++r) {
    size_t i = base + (size_t)r * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: element order inside a tile is (round, thread); ranks among
// equal digits come from 8 ballots per wave (lanes with the same digit),
// per-wave digit counts in LDS, and the running count of earlier rounds.
__global__ void __launch_bounds__(RS_THREADS) k_radix_scatter(const uint32_t* __restrict__ kin,
                                                               const uint32_t* __restrict__ vin, size_t n, int shift,
                                                               const uint32_t* __restrict__ offs, uint32_t nblocks,
                                                               uint32_t* __restrict__ kout,
                                                               uint32_t* __restrict__ vout) {
  __shared__ uint32_t running[256];
  __shared__ uint32_t wcnt[RS_THREADS / 64][256];
  const int wave = threadIdx.x >> 6;
  running[threadIdx.x] = offs[(size_t)threadIdx.x * nblocks + blockIdx.x];
  for (int w = 0; w < RS_THREADS / 64; ++w) wcnt[w][threadIdx.x] = 0;
  __syncthreads();
  size_t base = (size_t)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ROUNDS; ++r) {
    size_t i = base + (size_t)r * RS_THREADS + threadIdx.x;
    bool valid = i < n;
    uint32_t k = valid ? kin[i] : 0u;
    uint32_t v = valid ? vin[i] : 0u;
    uint32_t d = (k >> shift) & 255u;
    uint64_t match = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      uint64_t m = ballot(((d >> b) & 1u) != 0u);
      match &= ((d >> b) & 1u) ? m : ~m;
    }
    uint32_t rank = (uint32_t)__popcll(match & lanemask_lt());
    uint32_t cnt = (uint32_t)__popcll(match);
    bool leader = valid && rank == 0;
    if (leader) wcnt[wave][d] = cnt;
    __syncthreads();
    uint32_t pos = 0;
    if (valid) {
      pos = running[d] + rank;
      for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < RS_THREADS / 64; ++w) {
      add += wcnt[w][threadIdx.x];
      wcnt[w][threadIdx.x] = 0;
    }
    running[threadIdx.x] += add;
    __syncthreads();
  }
}

size_t radix_temp_words(size_t n) {
  size_t nb = (n + RS_TILE - 1) / RS_TILE;
  return 2 * 256 * nb + scan_temp_words(256 * nb) + 2 * n;
}

// Sorts (keys, vals) in place by key, stable; temp must hold radix_temp_words(n).
hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, size_t n, uint32_t* temp, hipStream_t st) {
  if (n <= 1) return hipSuccess;
  size_t nb = (n + RS_TILE - 1) / RS_TILE;
  uint32_t* hist = temp;
  uint32_t* offs = hist + 256 * nb;
  uint32_t* stmp = offs + 256 * nb;
  uint32_t* k2 = stmp + scan_temp_words(256 * nb);
  uint32_t* v2 = k2 + n;
  uint32_t *ka = keys, *va = vals, *kb = k2, *vb = v2;
  for (int pass = 0; pass < 4; ++pass) {
    int shift = pass * 8;
    hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)nb), dim3(RS_THREADS), 0, st, ka, n, shift, hist, (uint32_t)nb);
    hipError_t e = exclusive_scan(hist, offs, 256 * nb, stmp, nullptr, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)nb), dim3(RS_THREADS), 0, st, ka, va, n, shift, offs,
                       (uint32_t)nb, kb, vb);
    uint32_t* t = ka; ka = kb; kb = t;
    t = va; va = vb; vb = t;
  }
  // Four passes: the result is back in (keys, vals).
  return hipGetLastError();
}

hipError_t radix_pass(const uint32_t* kin, const uint32_t* vin, size_t n, int shift, uint32_t* temp, uint32_t* kout,
                      uint32_t* vout, hipStream_t st) {
  if (n == 0) return hipSuccess;
  size_t nb = (n + RS_TILE - 1) / RS_TILE;
  uint32_t* hist = temp;
  uint32_t* offs = hist + 256 * nb;
  uint32_t* stmp = offs + 256 * nb;
  hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)nb), dim3(RS_THREADS), 0, st, kin, n, shift, hist, (uint32_t)nb);
  hipError_t e = exclusive_scan(hist, offs, 256 * nb, stmp, nullptr, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)nb), dim3(RS_THREADS), 0, st, kin, vin, n, shift, offs,
                     (uint32_t)nb, kout, vout);
  return hipGetLastError();
}

// ---------------------------------------------------------------- fill
// A kernel rather than hipMemsetAsync, so that the stream work stays
// graph-capturable as kernel nodes only.
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t value, size_t n) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  size_t stride = (size_t)gridDim.x * 256;
  for (; i < n; i += stride) p[i] = value;
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  size_t nb = (n + 255) / 256;
  if (nb > 8192) nb = 8192;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)nb), dim3(256), 0, st, p, value, n);
  return hipGetLastError();
}

__global__ void k_gather_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ idx, uint32_t n,
                             uint32_t* __restrict__ dst) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

hipError_t launch_gather_u32(const uint32_t* src, const uint32_t* idx, uint32_t n, uint32_t* dst, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_u32, dim3((n + 255) / 256), dim3(256), 0, st, src, idx, n, dst);
  return hipGetLastError();
}

// ---------------------------------------------------------------- copy
__global__ void __launch_bounds__(256) k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  size_t stride = (size_t)gridDim.x * 256;
  for (; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_copy16(const void* src, void* dst, size_t bytes, hipStream_t st) {
  size_t n = bytes / 16;
  if (n == 0) return hipSuccess;
  size_t nb = (n + 255) / 256;
  if (nb > 8192) nb = 8192;
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)nb), dim3(256), 0, st, (const uint4*)src, (uint4*)dst, n);
  return hipGetLastError();
}

}  // namespace pm
