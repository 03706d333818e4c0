// pm_wave.h -- wave64 primitives for gfx950 (CDNA4): lane id, ballots,
// cross-lane moves and 64-bit key compare-exchange used by the sort, scan and
// kNN kernels.  Device-only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pm {

__device__ __forceinline__ int lane_id() {
  return (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ float uniform_f(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ uint32_t read_lane(uint32_t v, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}

// Cross-lane reads through the LDS crossbar (ds_bpermute; no LDS storage).
__device__ __forceinline__ uint32_t lane_read(uint32_t v, int src_lane) {
  return (uint32_t)__builtin_amdgcn_ds_bpermute(src_lane << 2, (int)v);
}
__device__ __forceinline__ float lane_read_f(float v, int src_lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(src_lane << 2, __builtin_bit_cast(int, v)));
}

// Lanes below this one in a 64-bit ballot mask.
__device__ __forceinline__ uint64_t lanemask_lt() {
  int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ bool key_less(uint32_t ah, uint32_t al, uint32_t bh, uint32_t bl) {
  return ah < bh || (ah == bh && al < bl);
}

// Bitonic sort of one 64-bit key per lane, ascending across the 64 lanes.
__device__ __forceinline__ void wave_bitonic_sort(uint32_t& h, uint32_t& l) {
  const int lane = lane_id();
#pragma unroll
  for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      uint32_t ph = lane_read(h, lane ^ stride);
      uint32_t pl = lane_read(l, lane ^ stride);
      bool asc = (lane & size) == 0 || size == 64;
      bool lower = (lane & stride) == 0;
      bool pless = key_less(ph, pl, h, l);
      bool take = (lower == asc) ? pless : key_less(h, l, ph, pl);
      if (take) {
        h = ph;
        l = pl;
      }
    }
  }
}

// Half-cleaner cascade: sorts a bitonic sequence held one key per lane.
__device__ __forceinline__ void wave_bitonic_clean(uint32_t& h, uint32_t& l) {
  const int lane = lane_id();
#pragma unroll
  for (int stride = 32; stride > 0; stride >>= 1) {
    uint32_t ph = lane_read(h, lane ^ stride);
    uint32_t pl = lane_read(l, lane ^ stride);
    bool lower = (lane & stride) == 0;
    bool take = lower ? key_less(ph, pl, h, l) : key_less(h, l, ph, pl);
    if (take) {
      h = ph;
      l = pl;
    }
  }
}

// Knuth's TwoSum: s + e == a + b exactly.
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Compensated butterfly sum over the wave; every lane ends with the same value.
__device__ __forceinline__ float wave_sum_comp(float v) {
  const int lane = lane_id();
  float s = v, c = 0.0f;
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
    float ps = lane_read_f(s, lane ^ m);
    float pc = lane_read_f(c, lane ^ m);
    float t, e;
    two_sum(s, ps, t, e);
    c = (c + pc) + e;
    s = t;
  }
  return s + c;
}

}  // namespace pm
