// pm_trace.hip -- photon emission and tracing on gfx950, one lane per photon.
//
// Replaces emit_photons_kernel / emitPhotons (photonMappingKernel.cu:1464-1480,
// :1215-1375) and init_photons_kernel (:1503-1521).  The reference splats
// energy into a shared voxel grid with racy read-modify-writes
// (:1068,1158,1177); here every photon writes its own records into slots
// fixed by a prefix sum over per-photon record counts, so the record array is
// in (photon id, event) order and bit-identical run to run:
//   pass 1  k_trace_count  : trace, count surface / volume records per photon
//   scan                   : exclusive prefix sums -> record offsets
//   pass 2  k_trace_write  : trace again, write records (SoA) at the offsets,
//                            and reduce the FP64 flux tally in a fixed order.
// Tracing twice costs far less than the gather it feeds and needs no
// per-photon staging buffers.
#include "pm_kernels.h"
#include "pm_wave.h"

namespace pm {

static constexpr int TRACE_THREADS = 256;

struct CountSink {
  uint32_t s, v;
  __device__ void surface(V3f, V3f, V3f) { ++s; }
  __device__ void volume(V3f, V3f, V3f) { ++v; }
};

__global__ void __launch_bounds__(TRACE_THREADS) k_trace_count(Scene sc, TraceParams tp, uint32_t pid0,
                                                                uint32_t npid, uint32_t* __restrict__ cnt_s,
                                                                uint32_t* __restrict__ cnt_v) {
  uint32_t i = blockIdx.x * TRACE_THREADS + threadIdx.x;
  if (i >= npid) return;
  CountSink sk{0u, 0u};
  Tally9 t;
  tally_zero(t);
  trace_photon(sc, tp, pid0 + i, sk, t);
  cnt_s[pid0 + i] = sk.s;
  cnt_v[pid0 + i] = sk.v;
}

__device__ __forceinline__ void write_record(const DevMap& m, uint32_t i, V3f p, V3f pw, V3f d) {
  m.x[i] = p.x;
  m.y[i] = p.y;
  m.z[i] = p.z;
  if (m.bf16) {
    uint4 w;
    w.x = (uint32_t)f32_to_bf16(pw.x) | ((uint32_t)f32_to_bf16(pw.y) << 16);
    w.y = (uint32_t)f32_to_bf16(pw.z) | ((uint32_t)f32_to_bf16(d.x) << 16);
    w.z = (uint32_t)f32_to_bf16(d.y) | ((uint32_t)f32_to_bf16(d.z) << 16);
    w.w = 0u;
    ((uint4*)m.payload)[i] = w;
  } else {
    float4* q = (float4*)m.payload;
    q[2 * (size_t)i] = make_float4(pw.x, pw.y, pw.z, d.x);
    q[2 * (size_t)i + 1] = make_float4(d.y, d.z, 0.0f, 0.0f);
  }
}

struct WriteSink {
  const DevMap* ms;
  const DevMap* mv;
  uint32_t is, iv;  // next record index of this photon in each map
  __device__ void surface(V3f p, V3f pw, V3f d) {
    if (is < ms->n) write_record(*ms, is, p, pw, d);
    ++is;
  }
  __device__ void volume(V3f p, V3f pw, V3f d) {
    if (iv < mv->n) write_record(*mv, iv, p, pw, d);
    ++iv;
  }
};

__global__ void __launch_bounds__(TRACE_THREADS) k_trace_write(Scene sc, TraceParams tp, uint32_t npid,
                                                                const uint32_t* __restrict__ off_s,
                                                                const uint32_t* __restrict__ off_v, DevMap ms,
                                                                DevMap mv, double* __restrict__ tally_partial) {
  __shared__ double red[TRACE_THREADS];
  uint32_t i = blockIdx.x * TRACE_THREADS + threadIdx.x;
  Tally9 t;
  tally_zero(t);
  if (i < npid) {
    WriteSink sk{&ms, &mv, off_s[i], off_v[i]};
    trace_photon(sc, tp, i, sk, t);
  }
  // Fixed-order tree reduction of each tally component over the block.
  for (int c = 0; c < 9; ++c) {
    red[threadIdx.x] = t.v[c];
    __syncthreads();
    for (int s = TRACE_THREADS / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) tally_partial[(size_t)blockIdx.x * 9 + c] = red[0];
    __syncthreads();
  }
}

// Batched emission: block b traces photons of scene block_scene[b], with the
// same per-photon work and the same per-block tally reduction as the
// one-scene kernels (record offsets come from one scan over all scenes).
__global__ void __launch_bounds__(TRACE_THREADS) k_trace_count_batch(const BatchTraceScene* __restrict__ scenes,
                                                                      const uint32_t* __restrict__ block_scene,
                                                                      uint32_t* __restrict__ cnt_s,
                                                                      uint32_t* __restrict__ cnt_v) {
  const BatchTraceScene& B = scenes[block_scene[blockIdx.x]];
  const uint32_t local = (blockIdx.x - B.block0) * TRACE_THREADS + threadIdx.x;
  const size_t g = (size_t)blockIdx.x * TRACE_THREADS + threadIdx.x;
  CountSink sk{0u, 0u};
  if (local < B.n_emit) {
    const Scene sc = B.sc;
    const TraceParams tp = B.tp;
    Tally9 t;
    tally_zero(t);
    trace_photon(sc, tp, local, sk, t);
  }
  cnt_s[g] = sk.s;
  cnt_v[g] = sk.v;
}

__global__ void __launch_bounds__(TRACE_THREADS) k_trace_write_batch(const BatchTraceScene* __restrict__ scenes,
                                                                      const uint32_t* __restrict__ block_scene,
                                                                      const uint32_t* __restrict__ off_s,
                                                                      const uint32_t* __restrict__ off_v, DevMap ms,
                                                                      DevMap mv, double* __restrict__ tally_partial) {
  __shared__ double red[TRACE_THREADS];
  const BatchTraceScene& B = scenes[block_scene[blockIdx.x]];
  const uint32_t local = (blockIdx.x - B.block0) * TRACE_THREADS + threadIdx.x;
  const size_t g = (size_t)blockIdx.x * TRACE_THREADS + threadIdx.x;
  Tally9 t;
  tally_zero(t);
  if (local < B.n_emit) {
    const Scene sc = B.sc;
    const TraceParams tp = B.tp;
    WriteSink sk{&ms, &mv, off_s[g], off_v[g]};
    trace_photon(sc, tp, local, sk, t);
  }
  for (int c = 0; c < 9; ++c) {
    red[threadIdx.x] = t.v[c];
    __syncthreads();
    for (int s = TRACE_THREADS / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) tally_partial[(size_t)blockIdx.x * 9 + c] = red[0];
    __syncthreads();
  }
}

// Sum of the per-block tallies, fixed order: thread t owns blocks t, t+256, ...
__global__ void __launch_bounds__(256) k_tally_final(const double* __restrict__ partial, uint32_t nblocks,
                                                      double* __restrict__ out9) {
  __shared__ double red[256];
  for (int c = 0; c < 9; ++c) {
    double s = 0.0;
    for (uint32_t b = threadIdx.x; b < nblocks; b += 256) s += partial[(size_t)b * 9 + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) out9[c] = red[0];
    __syncthreads();
  }
}

// One workgroup per scene: k_tally_final over the scene's own blocks.
__global__ void __launch_bounds__(256) k_tally_final_batch(const double* __restrict__ partial,
                                                            const BatchTraceScene* __restrict__ scenes,
                                                            double* __restrict__ out9) {
  __shared__ double red[256];
  const uint32_t b0 = scenes[blockIdx.x].block0, nb = scenes[blockIdx.x].nblocks;
  for (int c = 0; c < 9; ++c) {
    double s = 0.0;
    for (uint32_t b = threadIdx.x; b < nb; b += 256) s += partial[(size_t)(b0 + b) * 9 + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) out9[(size_t)blockIdx.x * 9 + c] = red[0];
    __syncthreads();
  }
}

// out = 1 + the photon id whose records reach `cap` (first p with
// off[p] < cap <= off[p] + cnt[p]); untouched if the total stays below cap.
__global__ void k_find_cap(const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt, uint32_t n,
                           uint32_t cap, uint32_t* __restrict__ out) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint64_t a = off[p], b = a + cnt[p];
  // one photon's range holds the cap-th record; the minimum keeps the
  // answer defined even if several ranges matched
  if (a < cap && b >= cap) atomicMin(out, p + 1u);
}

hipError_t launch_trace_count(const Scene& sc, const TraceParams& tp, uint32_t pid0, uint32_t npid, uint32_t* cnt_s,
                              uint32_t* cnt_v, hipStream_t st) {
  if (npid == 0) return hipSuccess;
  unsigned nb = (npid + TRACE_THREADS - 1) / TRACE_THREADS;
  hipLaunchKernelGGL(k_trace_count, dim3(nb), dim3(TRACE_THREADS), 0, st, sc, tp, pid0, npid, cnt_s, cnt_v);
  return hipGetLastError();
}

uint32_t trace_write_blocks(uint32_t npid) { return (npid + TRACE_THREADS - 1) / TRACE_THREADS; }

hipError_t launch_trace_write(const Scene& sc, const TraceParams& tp, uint32_t npid, const uint32_t* off_s,
                              const uint32_t* off_v, DevMap ms, DevMap mv, double* tally_partial, hipStream_t st) {
  if (npid == 0) return hipSuccess;
  unsigned nb = trace_write_blocks(npid);
  hipLaunchKernelGGL(k_trace_write, dim3(nb), dim3(TRACE_THREADS), 0, st, sc, tp, npid, off_s, off_v, ms, mv,
                     tally_partial);
  return hipGetLastError();
}

hipError_t launch_tally_final(const double* partial, uint32_t nblocks, double* out9, hipStream_t st) {
  hipLaunchKernelGGL(k_tally_final, dim3(1), dim3(256), 0, st, partial, nblocks, out9);
  return hipGetLastError();
}

hipError_t launch_find_cap(const uint32_t* off, const uint32_t* cnt, uint32_t n, uint32_t cap, uint32_t* out,
                           hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_find_cap, dim3((n + 255) / 256), dim3(256), 0, st, off, cnt, n, cap, out);
  return hipGetLastError();
}

uint32_t trace_threads() { return TRACE_THREADS; }

hipError_t launch_trace_count_batch(const BatchTraceScene* scenes, const uint32_t* block_scene, uint32_t nblocks,
                                    uint32_t* cnt_s, uint32_t* cnt_v, hipStream_t st) {
  if (nblocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_trace_count_batch, dim3(nblocks), dim3(TRACE_THREADS), 0, st, scenes, block_scene, cnt_s,
                     cnt_v);
  return hipGetLastError();
}

hipError_t launch_trace_write_batch(const BatchTraceScene* scenes, const uint32_t* block_scene, uint32_t nblocks,
                                    const uint32_t* off_s, const uint32_t* off_v, DevMap ms, DevMap mv,
                                    double* tally_partial, hipStream_t st) {
  if (nblocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_trace_write_batch, dim3(nblocks), dim3(TRACE_THREADS), 0, st, scenes, block_scene, off_s,
                     off_v, ms, mv, tally_partial);
  return hipGetLastError();
}

hipError_t launch_tally_final_batch(const double* partial, const BatchTraceScene* scenes, uint32_t nscenes,
                                    double* out9, hipStream_t st) {
  if (nscenes == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tally_final_batch, dim3(nscenes), dim3(256), 0, st, partial, scenes, out9);
  return hipGetLastError();
}

}  // namespace pm
