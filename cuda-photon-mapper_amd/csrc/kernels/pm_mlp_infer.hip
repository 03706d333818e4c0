// pm_mlp_infer.hip -- fused inference of the radiance-cache MLP on gfx950.
//
// The generic GEMM in pm_mlp.hip stages every operand through LDS per tile,
// which suits training but wastes bandwidth for inference over hundreds of
// millions of query rows.  Inference therefore runs as two persistent kernels
// whose weights stay resident in LDS for the whole launch:
//   k_mlp_l1  : encoding computed on the fly as MFMA A-fragments (two
//               sin/cos pairs per lane per 16-wide K step) x W1 (LDS, 59 KB)
//               -> bias, ReLU -> H1 (BF16, row-major)
//   k_mlp_l25 : H1 fragments -> W2, W3, W4 (LDS) with bias + ReLU, the
//               activations passing through a per-wave LDS stage between
//               layers -> W5 -> radiance = max(0, exp(y) - 1)
// Each wave handles 32 rows (two 16-row MFMA groups) per step so every
// B-fragment read from LDS feeds two v_mfma_f32_16x16x16_bf16.
// Lane layout of that instruction: A[i][k] in lane i + 16*(k/4), element k%4;
// B[k][j] in lane j + 16*(k/4); D[i][j] in lane j + 16*(i/4), element i%4.
#include "pm_kernels.h"
#include "pm_mlp.h"
#include "pm_wave.h"

namespace pm {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

static constexpr int IW = 256;                 // threads per workgroup (4 waves)
static constexpr int W1S = MLP_KPAD + 8;       // LDS row stride of W1 (bf16)
static constexpr int WHS = MLP_HID + 8;        // LDS row stride of W2..W5 and the stage
static constexpr int ROWS_PER_WAVE = 32;

__device__ __forceinline__ bf16x4 lds_frag(const uint16_t* p) {
  uint2 v = *(const uint2*)p;
  bf16x4 r;
  r[0] = (short)(v.x & 0xffffu);
  r[1] = (short)(v.x >> 16);
  r[2] = (short)(v.y & 0xffffu);
  r[3] = (short)(v.y >> 16);
  return r;
}

// ---------------------------------------------------------------- layer 1
__global__ void __launch_bounds__(IW) k_mlp_l1(const float* __restrict__ in, uint32_t rows,
                                                const uint16_t* __restrict__ w1, const float* __restrict__ b1,
                                                uint16_t* __restrict__ h1) {
  __shared__ __align__(16) uint16_t W[MLP_HID * W1S];
  // 16-byte copies: rows of 224 bf16 = 28 x 8, LDS row stride 464 B
  for (int e = threadIdx.x; e < MLP_HID * (MLP_KPAD / 8); e += IW) {
    int n = e / (MLP_KPAD / 8), k8 = e % (MLP_KPAD / 8);
    *(uint4*)&W[n * W1S + 8 * k8] = ((const uint4*)w1)[e];
  }
  __syncthreads();
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  const int r16 = lane & 15, q = lane >> 4;
  const uint32_t tiles = (rows + ROWS_PER_WAVE - 1) / ROWS_PER_WAVE;
  for (uint32_t t = blockIdx.x * (IW / 64) + wave; t < tiles; t += gridDim.x * (IW / 64)) {
    const uint32_t row0 = t * ROWS_PER_WAVE;
    f32x4 acc[2][8];
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int c = 0; c < 8; ++c) acc[g][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int ks = 0; ks < MLP_KPAD / 16; ++ks) {
      const int f0 = ks * 16 + 4 * q;  // first of this lane's 4 features
      bf16x4 a[2];
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        const uint32_t row = row0 + g * 16 + r16;
        float s0 = 0.0f, c0 = 0.0f, s1 = 0.0f, c1 = 0.0f;
        if (f0 < MLP_FEAT && row < rows) {
          const int i = f0 / (2 * MLP_BANDS), b = (f0 % (2 * MLP_BANDS)) / 2;
          const float v = in[(size_t)row * MLP_IN + i];
          const float sc = u2f((uint32_t)(126 + b) << 23);  // 2^(b-1)
          sincos_turns(v * sc, &s0, &c0);
          sincos_turns(v * (sc * 2.0f), &s1, &c1);
        }
        a[g] = bf16x4{(short)f32_to_bf16(s0), (short)f32_to_bf16(c0), (short)f32_to_bf16(s1),
                      (short)f32_to_bf16(c1)};
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        bf16x4 bb = lds_frag(&W[(c * 16 + r16) * W1S + ks * 16 + 4 * q]);
        acc[0][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0], bb, acc[0][c], 0, 0, 0);
        acc[1][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[1], bb, acc[1][c], 0, 0, 0);
      }
    }
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int col = c * 16 + r16;
        const float bias = b1[col];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const uint32_t row = row0 + g * 16 + 4 * q + e;
          if (row < rows) {
            float v = acc[g][c][e] + bias;
            h1[(size_t)row * MLP_HID + col] = f32_to_bf16(v > 0.0f ? v : 0.0f);
          }
        }
      }
  }
}

// ---------------------------------------------------------------- layers 2..5
struct L25Args {
  const uint16_t* h1;
  uint32_t rows;
  const uint16_t* w[4];  // W2, W3, W4 [128 x 128], W5 [16 x 128]
  const float* b[4];
  float* rgb;            // [rows x 3]
};

__global__ void __launch_bounds__(IW) k_mlp_l25(L25Args A) {
  extern __shared__ uint16_t lds[];
  uint16_t* W = lds;                                  // 3 x 128 x WHS, then 16 x WHS
  uint16_t* stage_all = lds + (3 * MLP_HID + MLP_OUTPAD) * WHS;
  for (int l = 0; l < 4; ++l) {
    const int nout = l == 3 ? MLP_OUTPAD : MLP_HID;
    // 16-byte copies: rows of 128 bf16 = 16 x 8, LDS row stride 272 B
    for (int e = threadIdx.x; e < nout * (MLP_HID / 8); e += IW) {
      int n = e / (MLP_HID / 8), k8 = e % (MLP_HID / 8);
      *(uint4*)&W[(l * MLP_HID + n) * WHS + 8 * k8] = ((const uint4*)A.w[l])[e];
    }
  }
  __syncthreads();
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  const int r16 = lane & 15, q = lane >> 4;
  uint16_t* stage = stage_all + wave * ROWS_PER_WAVE * WHS;
  const uint32_t tiles = (A.rows + ROWS_PER_WAVE - 1) / ROWS_PER_WAVE;
  for (uint32_t t = blockIdx.x * (IW / 64) + wave; t < tiles; t += gridDim.x * (IW / 64)) {
    const uint32_t row0 = t * ROWS_PER_WAVE;
    // A fragments of H1 straight from global memory (8 bytes per lane per K step)
    bf16x4 a[2][8];
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const uint32_t row = row0 + g * 16 + r16;
#pragma unroll
      for (int ks = 0; ks < 8; ++ks) {
        if (row < A.rows) {
          uint2 v = *(const uint2*)&A.h1[(size_t)row * MLP_HID + ks * 16 + 4 * q];
          a[g][ks] = bf16x4{(short)(v.x & 0xffffu), (short)(v.x >> 16), (short)(v.y & 0xffffu), (short)(v.y >> 16)};
        } else {
          a[g][ks] = bf16x4{0, 0, 0, 0};
        }
      }
    }
    for (int l = 0; l < 3; ++l) {
      const uint16_t* Wl = W + l * MLP_HID * WHS;
      f32x4 acc[2][8];
#pragma unroll
      for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[g][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < 8; ++ks)
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          bf16x4 bb = lds_frag(&Wl[(c * 16 + r16) * WHS + ks * 16 + 4 * q]);
          acc[0][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0][ks], bb, acc[0][c], 0, 0, 0);
          acc[1][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[1][ks], bb, acc[1][c], 0, 0, 0);
        }
      // bias + ReLU, D layout -> stage (row-major) -> A layout for the next layer
#pragma unroll
      for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const int col = c * 16 + r16;
          const float bias = A.b[l][col];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            float v = acc[g][c][e] + bias;
            stage[(g * 16 + 4 * q + e) * WHS + col] = f32_to_bf16(v > 0.0f ? v : 0.0f);
          }
        }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#pragma unroll
      for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) a[g][ks] = lds_frag(&stage[(g * 16 + r16) * WHS + ks * 16 + 4 * q]);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    // output layer: 16 (padded) columns, first 3 used
    const uint16_t* W5 = W + 3 * MLP_HID * WHS;
    f32x4 y[2] = {f32x4{0.0f, 0.0f, 0.0f, 0.0f}, f32x4{0.0f, 0.0f, 0.0f, 0.0f}};
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      bf16x4 bb = lds_frag(&W5[r16 * WHS + ks * 16 + 4 * q]);
      y[0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0][ks], bb, y[0], 0, 0, 0);
      y[1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[1][ks], bb, y[1], 0, 0, 0);
    }
    if (r16 < MLP_OUT) {
      const float bias = A.b[3][r16];
#pragma unroll
      for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const uint32_t row = row0 + g * 16 + 4 * q + e;
          if (row < A.rows) {
            float v = exp_(y[g][e] + bias) - 1.0f;
            A.rgb[(size_t)row * MLP_OUT + r16] = v > 0.0f ? v : 0.0f;
          }
        }
    }
  }
}

// ---------------------------------------------------------------- K = 32 variants
// gfx950's v_mfma_f32_16x16x32_bf16 does twice the work of the 16x16x16
// instruction per issue.  Lane layout: A[i][k] in lane i + 16*(k/8), element
// k%8; B[k][j] in lane j + 16*(k/8); D as for 16x16x16.  A lane's 8 features of
// one K step are 4 consecutive bands (sin, cos) of one input, so layer 1
// computes one sin/cos and derives the next 3 bands by the double-angle
// identities (sin 2x = 2 sin x cos x, cos 2x = (cos x - sin x)(cos x + sin x)),
// whose FP32 error is far below the BF16 rounding of the features.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ bf16x8 lds_frag8(const uint16_t* p) {
  uint4 v = *(const uint4*)p;
  return *(const bf16x8*)&v;
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// Two FP32 values rounded to BF16 (nearest even) in one v_cvt_pk_bf16_f32:
// a in the low half.  For finite inputs it equals f32_to_bf16 on each.
__device__ __forceinline__ uint32_t cvt_pk_bf16(float a, float b) {
  const bf16x2 r = __builtin_convertvector((f32x2){a, b}, bf16x2);
  return *(const uint32_t*)&r;
}

__device__ __forceinline__ bf16x8 pack8(const float* f) {
  uint4 v;
  v.x = cvt_pk_bf16(f[0], f[1]);
  v.y = cvt_pk_bf16(f[2], f[3]);
  v.z = cvt_pk_bf16(f[4], f[5]);
  v.w = cvt_pk_bf16(f[6], f[7]);
  return *(const bf16x8*)&v;
}

// Bias + ReLU of a D fragment's 4 rows of one column, written to the
// row-major stage as BF16 (two packed conversions, four 2-byte stores).
__device__ __forceinline__ void stage_col4(uint16_t* stage, int stride, int q, int col, f32x4 acc, float bias) {
  float v[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const float x = acc[e] + bias;
    v[e] = x > 0.0f ? x : 0.0f;
  }
  const uint32_t p01 = cvt_pk_bf16(v[0], v[1]), p23 = cvt_pk_bf16(v[2], v[3]);
  stage[(4 * q + 0) * stride + col] = (uint16_t)(p01 & 0xffffu);
  stage[(4 * q + 1) * stride + col] = (uint16_t)(p01 >> 16);
  stage[(4 * q + 2) * stride + col] = (uint16_t)(p23 & 0xffffu);
  stage[(4 * q + 3) * stride + col] = (uint16_t)(p23 >> 16);
}

static constexpr int STS = MLP_HID + 8;  // stage row stride (bf16), 272 B

// Layer 1: features on the fly, W1 in LDS, H1 written through a per-wave
// 16-row stage as 16-byte row segments.  12 waves share W1 (3 per SIMD).
static constexpr int IW1X = 768;

__global__ void __launch_bounds__(IW1X) k_mlp_l1x(const float* __restrict__ in, uint32_t rows,
                                                   const uint16_t* __restrict__ w1, const float* __restrict__ b1,
                                                   uint16_t* __restrict__ h1) {
  __shared__ __align__(16) uint16_t W[MLP_HID * W1S];
  __shared__ __align__(16) uint16_t stage_all[(IW1X / 64) * 16 * STS];
  for (int e = threadIdx.x; e < MLP_HID * (MLP_KPAD / 8); e += IW1X) {
    int n = e / (MLP_KPAD / 8), k8 = e % (MLP_KPAD / 8);
    *(uint4*)&W[n * W1S + 8 * k8] = ((const uint4*)w1)[e];
  }
  __syncthreads();
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  const int r16 = lane & 15, q = lane >> 4;
  uint16_t* stage = stage_all + wave * 16 * STS;
  const uint32_t tiles = (rows + ROWS_PER_WAVE - 1) / ROWS_PER_WAVE;
  for (uint32_t t = blockIdx.x * (IW1X / 64) + wave; t < tiles; t += gridDim.x * (IW1X / 64)) {
    const uint32_t row0 = t * ROWS_PER_WAVE;
    f32x4 acc[2][8];
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int c = 0; c < 8; ++c) acc[g][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int ks = 0; ks < MLP_KPAD / 32; ++ks) {
      const int f0 = ks * 32 + 8 * q;  // this lane's 8 features: 4 bands of one input
      bf16x8 a[2];
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        const uint32_t row = row0 + g * 16 + r16;
        float f[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (f0 < MLP_FEAT && row < rows) {
          const int i = f0 / (2 * MLP_BANDS), b = (f0 % (2 * MLP_BANDS)) / 2;
          const float v = in[(size_t)row * MLP_IN + i];
          float sn, cs;
          sincos_turns(v * u2f((uint32_t)(126 + b) << 23), &sn, &cs);  // 2^(b-1)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            f[2 * j] = sn;
            f[2 * j + 1] = cs;
            const float s2 = 2.0f * (sn * cs), c2 = (cs - sn) * (cs + sn);
            sn = s2;
            cs = c2;
          }
        }
        a[g] = pack8(f);
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        bf16x8 bb = lds_frag8(&W[(c * 16 + r16) * W1S + ks * 32 + 8 * q]);
        acc[0][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bb, acc[0][c], 0, 0, 0);
        acc[1][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], bb, acc[1][c], 0, 0, 0);
      }
    }
#pragma unroll
    for (int g = 0; g < 2; ++g) {
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int col = c * 16 + r16;
        stage_col4(stage, STS, q, col, acc[g][c], b1[col]);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      // 16 rows x 256 B = 256 segments of 16 B: 4 per lane
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        const int seg = s4 * 64 + lane, r = seg >> 4, c8 = seg & 15;
        const uint32_t row = row0 + g * 16 + r;
        if (row < rows) *(uint4*)&h1[(size_t)row * MLP_HID + 8 * c8] = *(const uint4*)&stage[r * STS + 8 * c8];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
  }
}

// Layers 2..5 with K = 32 steps: 12 waves share the weights (3 per SIMD; the
// weights and the 12 stages take 157 KB of the 160 KB of LDS); each wave
// converts its 32-row activations D -> A through a 16-row LDS stage.
static constexpr int IW25 = 768;

__global__ void __launch_bounds__(IW25) k_mlp_l25x(L25Args A) {
  extern __shared__ uint16_t lds[];
  uint16_t* W = lds;
  uint16_t* stage_all = lds + (3 * MLP_HID + MLP_OUTPAD) * WHS;
  for (int l = 0; l < 4; ++l) {
    const int nout = l == 3 ? MLP_OUTPAD : MLP_HID;
    for (int e = threadIdx.x; e < nout * (MLP_HID / 8); e += IW25) {
      int n = e / (MLP_HID / 8), k8 = e % (MLP_HID / 8);
      *(uint4*)&W[(l * MLP_HID + n) * WHS + 8 * k8] = ((const uint4*)A.w[l])[e];
    }
  }
  __syncthreads();
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  const int r16 = lane & 15, q = lane >> 4;
  uint16_t* stage = stage_all + wave * 16 * WHS;
  const uint32_t tiles = (A.rows + ROWS_PER_WAVE - 1) / ROWS_PER_WAVE;
  for (uint32_t t = blockIdx.x * (IW25 / 64) + wave; t < tiles; t += gridDim.x * (IW25 / 64)) {
    const uint32_t row0 = t * ROWS_PER_WAVE;
    bf16x8 a[2][4];
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const uint32_t row = row0 + g * 16 + r16;
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (row < A.rows) v = *(const uint4*)&A.h1[(size_t)row * MLP_HID + ks * 32 + 8 * q];
        a[g][ks] = *(const bf16x8*)&v;
      }
    }
    for (int l = 0; l < 3; ++l) {
      const uint16_t* Wl = W + l * MLP_HID * WHS;
      f32x4 acc[2][8];
#pragma unroll
      for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[g][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks)
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          bf16x8 bb = lds_frag8(&Wl[(c * 16 + r16) * WHS + ks * 32 + 8 * q]);
          acc[0][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][ks], bb, acc[0][c], 0, 0, 0);
          acc[1][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][ks], bb, acc[1][c], 0, 0, 0);
        }
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        // bias + ReLU, D layout -> stage (row-major) -> A layout for the next layer
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const int col = c * 16 + r16;
          stage_col4(stage, WHS, q, col, acc[g][c], A.b[l][col]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) a[g][ks] = lds_frag8(&stage[r16 * WHS + ks * 32 + 8 * q]);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      }
    }
    const uint16_t* W5 = W + 3 * MLP_HID * WHS;
    f32x4 y[2] = {f32x4{0.0f, 0.0f, 0.0f, 0.0f}, f32x4{0.0f, 0.0f, 0.0f, 0.0f}};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      bf16x8 bb = lds_frag8(&W5[r16 * WHS + ks * 32 + 8 * q]);
      y[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][ks], bb, y[0], 0, 0, 0);
      y[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][ks], bb, y[1], 0, 0, 0);
    }
    if (r16 < MLP_OUT) {
      const float bias = A.b[3][r16];
#pragma unroll
      for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const uint32_t row = row0 + g * 16 + 4 * q + e;
          if (row < A.rows) {
            float v = exp_(y[g][e] + bias) - 1.0f;
            A.rgb[(size_t)row * MLP_OUT + r16] = v > 0.0f ? v : 0.0f;
          }
        }
    }
  }
}

// ---------------------------------------------------------------- one fused kernel
// All five layers in one kernel, the activations never leaving registers.
// The layers are computed transposed, out^T = W . H^T: the weights are the
// A operand (read from LDS) and a wave's 64 rows are the B operand, as four
// 16-row tiles.  The D fragment of m-tile c holds, in lane l, neurons
// 16c + 4(l/16) + e (e = 0..3) of row l%16, and that is almost the B layout
// the next layer needs (8 consecutive K values of row l%16 per lane).  The
// K order inside an MFMA is free as long as A and B agree, so K step s of the
// next layer takes, in lane group g = l/16, the neurons
//   n(s, 8g + e) = 32s + 16(e / 4) + 4g + e%4,   e = 0..7,
// i.e. the D fragments of m-tiles 2s and 2s + 1 of the same lane, and its
// weights are stored in LDS with their columns in that order.  No LDS stage,
// no cross-lane exchange.  Each weight fragment read from LDS feeds 4 MFMAs.
//
// LDS (bf16): every layer as K/32 blocks of [nout x 32] (64 B rows, four
// 16-B chunks; chunk g of row r at position g ^ ((r / 4) % 4), so the 8
// lanes of a read phase hit distinct banks): W1 7 x 128 rows, W2..W4 4 x 128,
// W5 4 x 16.  159,744 B, one workgroup of 8 waves (2 per SIMD) per CU.
static constexpr int IWF = 512;
static constexpr int FT = 4;                       // 16-row tiles per wave
static constexpr int FROWS = 16 * FT;              // rows per wave step
static constexpr uint32_t FW1 = 0;                 // byte offsets of the layers
static constexpr uint32_t FW2 = FW1 + 7u * 128u * 64u;
static constexpr uint32_t FW3 = FW2 + 4u * 128u * 64u;
static constexpr uint32_t FW4 = FW3 + 4u * 128u * 64u;
static constexpr uint32_t FW5 = FW4 + 4u * 128u * 64u;
static constexpr uint32_t FB14 = FW5 + 4u * 16u * 64u;  // biases of layers 1..4, 4 x 128 FP32
static constexpr uint32_t FLDS = FB14 + 4u * 128u * 4u;

__device__ __forceinline__ uint32_t fchunk(uint32_t nout, uint32_t ks, uint32_t r, uint32_t g) {
  return ks * nout * 64u + r * 64u + 16u * (g ^ ((r >> 2) & 3u));
}

// Layer 1's K order in the fused kernel: lane group g takes inputs 2g and
// 2g + 1 in K steps 0..5 (bands 0-3, 4-7, 8-11 of each), and K step 6 takes
// bands 4g .. 4g + 3 of input 8 (g < 3; g = 3 is padding).  A lane then needs
// three inputs per row and five sin/cos evaluations (bands 0 and 8 of two
// inputs, band 4g of input 8), the other bands by the double-angle identities.  Returns the first of the 8 encoding columns
// (24 input + 2 band + {sin, cos}) of step ks, group g, or -1 for padding.
__host__ __device__ __forceinline__ int l1_input(int ks, int g) { return ks < 6 ? 2 * g + ks / 3 : (g < 3 ? 8 : -1); }
__host__ __device__ __forceinline__ int l1_band0(int ks, int g) { return ks < 6 ? 4 * (ks % 3) : 4 * g; }
__host__ __device__ __forceinline__ int l1_feature_col(int ks, int g) {
  const int i = l1_input(ks, g);
  return i < 0 ? -1 : 2 * MLP_BANDS * i + 2 * l1_band0(ks, g);
}

// Volume records (normal 0) in 5 K steps: the features of inputs 6..8 are the
// constants sin 0 = 0 and cos 0 = 1, so their products with W1 are folded into
// layer 1's bias, and only the 144 features of position and direction remain.
// Lane group g takes input g in K steps 0..2 (bands 0-3, 4-7, 8-11), and in
// K steps 3..4 input 4 + (g & 1): groups 0 and 1 bands 0-3 then 4-7, groups
// 2 and 3 bands 8-11 (then padding).  Two inputs and four sin/cos evaluations
// per lane and row.
__host__ __device__ __forceinline__ int l1v_input(int ks, int g) {
  return ks < 3 ? g : (ks == 3 || g < 2 ? 4 + (g & 1) : -1);
}
__host__ __device__ __forceinline__ int l1v_band0(int ks, int g) { return ks < 3 ? 4 * ks : (ks == 3 ? (g < 2 ? 0 : 8) : 4); }
__host__ __device__ __forceinline__ int l1v_feature_col(int ks, int g) {
  const int i = l1v_input(ks, g);
  return i < 0 ? -1 : 2 * MLP_BANDS * i + 2 * l1v_band0(ks, g);
}

struct FusedArgs {
  const float* in;
  uint32_t rows;
  const uint32_t* list;  // row i of the launch is record list[i] (in and rgb); null: record i
  const uint16_t* w[MLP_LAYERS];  // BF16 weights, row-major [nout x kin]
  const float* b[MLP_LAYERS];
  float* rgb;
};

// ReLU of m-tiles 2s and 2s + 1 (their accumulators started from the bias)
// into the next layer's B fragment for K step s.  The ReLU is applied after
// the BF16 rounding, as a packed 16-bit integer max with 0: rounding keeps the
// sign, and a BF16 value is negative exactly when its bit pattern is as a
// signed integer (-0 becomes +0), so the bits equal those of rounding max(x, 0).
typedef short s16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t relu_pk(uint32_t p) {
  s16x2 v = *(const s16x2*)&p;
  v = __builtin_elementwise_max(v, (s16x2){0, 0});
  return *(const uint32_t*)&v;
}
__device__ __forceinline__ bf16x8 next_b(const f32x4& d0, const f32x4& d1) {
  uint4 v;
  v.x = relu_pk(cvt_pk_bf16(d0[0], d0[1]));
  v.y = relu_pk(cvt_pk_bf16(d0[2], d0[3]));
  v.z = relu_pk(cvt_pk_bf16(d1[0], d1[1]));
  v.w = relu_pk(cvt_pk_bf16(d1[2], d1[3]));
  return *(const bf16x8*)&v;
}

// VOL: every row is a volume record (normal 0), layer 1 in KS1 = 5 K steps.
template <bool VOL>
__global__ void __launch_bounds__(IWF) k_mlp_fused(FusedArgs A) {
  extern __shared__ __align__(16) unsigned char flds[];
  constexpr int KS1 = VOL ? 5 : MLP_KPAD / 32;  // layer 1's K steps
  // weights -> LDS: one 16-B chunk per step; layers 2..5 in the permuted K order
  {
    const uint32_t c1 = 128u * 4u * (uint32_t)KS1, c24 = 3u * 128u * 16u, c5 = 16u * 16u;
    for (uint32_t e = threadIdx.x; e < c1 + c24 + c5; e += IWF) {
      uint4 v;
      uint32_t dst;
      if (e < c1) {
        // W1 row r, K step ks, lane group g: 8 features (4 bands of one input)
        const uint32_t r = e / (4u * KS1), k8 = e % (4u * KS1), ks = k8 >> 2, g = k8 & 3u;
        const int col = VOL ? l1v_feature_col((int)ks, (int)g) : l1_feature_col((int)ks, (int)g);
        v = col < 0 ? make_uint4(0u, 0u, 0u, 0u) : *(const uint4*)&A.w[0][(size_t)r * MLP_KPAD + col];
        dst = FW1 + fchunk(128u, ks, r, g);
      } else {
        const uint32_t e2 = e - c1;
        const uint32_t l = e2 < c24 ? 1u + e2 / (128u * 16u) : 4u;
        const uint32_t f = e2 < c24 ? e2 % (128u * 16u) : e2 - c24;
        const uint32_t r = f / 16u, kc = f % 16u, ks = kc >> 2, g = kc & 3u;
        const uint16_t* w = A.w[l] + (size_t)r * MLP_HID;
        const uint2 lo = *(const uint2*)&w[32u * ks + 4u * g];
        const uint2 hi = *(const uint2*)&w[32u * ks + 16u + 4u * g];
        v = make_uint4(lo.x, lo.y, hi.x, hi.y);
        const uint32_t base = l == 1u ? FW2 : (l == 2u ? FW3 : (l == 3u ? FW4 : FW5));
        dst = base + fchunk(l == 4u ? 16u : 128u, ks, r, g);
      }
      *(uint4*)&flds[dst] = v;
    }
    float* bl = (float*)&flds[FB14];
    for (uint32_t e = threadIdx.x; e < 4u * 128u; e += IWF) {
      float b = A.b[e >> 7][e & 127u];
      if (VOL && e < 128u) {
        // + W1 x the constant features of inputs 6..8 (the cos columns, 1)
        const uint16_t* w = A.w[0] + (size_t)e * MLP_KPAD;
        float f = 0.0f;
        for (int c = 2 * MLP_BANDS * 6 + 1; c < MLP_FEAT; c += 2) f += bf16_to_f32(w[c]);
        b += f;
      }
      bl[e] = b;
    }
  }
  __syncthreads();
  const float* bias_l = (const float*)&flds[FB14];
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  const int j = lane & 15, g = lane >> 4;
  const uint32_t steps = (A.rows + FROWS - 1) / FROWS;
  // a lane's three inputs of each tile (zero past the end): loaded one wave
  // step ahead, so their latency hides behind the previous step's MFMAs
  float x0[FT], x1[FT], v8[FT];
  auto load_inputs = [&](uint32_t t) {
#pragma unroll
    for (int u = 0; u < FT; ++u) {
      const uint32_t row = t * FROWS + 16u * (uint32_t)u + (uint32_t)j;
      const bool ok = t < steps && row < A.rows;
      const uint32_t rec = !ok ? 0u : (A.list ? A.list[row] : row);
      const float* x = A.in + (size_t)rec * MLP_IN;
      if (VOL) {
        x0[u] = ok ? x[g] : 0.0f;
        x1[u] = ok ? x[4 + (g & 1)] : 0.0f;
      } else {
        x0[u] = ok ? x[2 * g] : 0.0f;
        x1[u] = ok ? x[2 * g + 1] : 0.0f;
        v8[u] = ok && g < 3 ? x[8] : 0.0f;
      }
    }
  };
  const uint32_t tstride = gridDim.x * (IWF / 64);
  for (uint32_t t = blockIdx.x * (IWF / 64) + wave; t < steps; t += tstride) {
    load_inputs(t);
    const uint32_t row0 = t * FROWS;
    // layer 1: the encoding computed as B fragments, in the K order of
    // l1_feature_col: per tile, a lane's three inputs (bands 0 and 8 of the
    // first two and band 4g of input 8 evaluated directly, the other bands by
    // the double-angle identities: at most 7 doublings in a row keep the
    // features within 1e-5 of direct values)
    float sn[FT], cs[FT];
    auto features = [&](int ks, bf16x8* bf) {
#pragma unroll
      for (int u = 0; u < FT; ++u) {
        if (ks == 0) sincos_turns(x0[u] * 0.5f, &sn[u], &cs[u]);    // band 0: 2^-1 turns per unit
        if (ks == 2) sincos_turns(x0[u] * 128.0f, &sn[u], &cs[u]);  // band 8: 2^7 turns per unit
        if (VOL) {
          if (ks == 3) sincos_turns(x1[u] * (g < 2 ? 0.5f : 128.0f), &sn[u], &cs[u]);  // band 0 or 8
        } else {
          if (ks == 3) sincos_turns(x1[u] * 0.5f, &sn[u], &cs[u]);  // the lane's second input
          if (ks == 5) sincos_turns(x1[u] * 128.0f, &sn[u], &cs[u]);
          if (ks == 6) sincos_turns(v8[u] * u2f((uint32_t)(126 + 4 * g) << 23), &sn[u], &cs[u]);  // band 4g
        }
        float f[8];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          f[2 * q] = sn[u];
          f[2 * q + 1] = cs[u];
          const float s2 = 2.0f * (sn[u] * cs[u]), c2 = (cs[u] - sn[u]) * (cs[u] + sn[u]);
          sn[u] = s2;
          cs[u] = c2;
        }
        // rows past the end (inputs 0) only feed their own outputs, which are
        // not written, and the padding columns of W1 are zero in LDS: no masking
        bf[u] = pack8(f);
      }
    };
    // the accumulators start from the bias of their neurons 16c + 4g + e
    auto bias4 = [&](const float* b, int c) { return *(const f32x4*)&b[16 * c + 4 * g]; };
    f32x4 acc[8][FT];
    {
      // K steps fully unrolled, the features of step ks + 1 computed while the
      // MFMAs of step ks run (two fragment buffers, no copies)
      bf16x8 fa[FT], fb[FT];
      features(0, fa);
#pragma unroll
      for (int ks = 0; ks < KS1; ++ks) {
        bf16x8* cur = (ks & 1) ? fb : fa;
        bf16x8* nxt = (ks & 1) ? fa : fb;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const bf16x8 a = *(const bf16x8*)&flds[FW1 + fchunk(128u, (uint32_t)ks, (uint32_t)(16 * c + j), (uint32_t)g)];
          const f32x4 b0 = ks == 0 ? bias4(bias_l, c) : f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
          for (int u = 0; u < FT; ++u)
            acc[c][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, cur[u], ks == 0 ? b0 : acc[c][u], 0, 0, 0);
        }
        if (ks + 1 < KS1) features(ks + 1, nxt);
        // keep each step's work in its step: scheduling the unrolled steps
        // together would raise the register pressure past the budget
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    bf16x8 hb[4][FT];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int u = 0; u < FT; ++u) hb[s][u] = next_b(acc[2 * s][u], acc[2 * s + 1][u]);
    // layers 2..4, one pair of m-tiles at a time: pair p completes the next
    // layer's K step p, whose ReLU + packing then overlaps the MFMAs of pair
    // p + 1 (and only two m-tiles of accumulators are live)
#pragma unroll
    for (int l = 1; l < MLP_HIDDEN_LAYERS; ++l) {
      const uint32_t base = l == 1 ? FW2 : (l == 2 ? FW3 : FW4);
      const float* bl = bias_l + 128 * l;
      bf16x8 nb[4][FT];
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        f32x4 d[2][FT];
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
          const f32x4 b0 = bias4(bl, 2 * p + cc);
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const bf16x8 a =
                *(const bf16x8*)&flds[base + fchunk(128u, (uint32_t)s, (uint32_t)(16 * (2 * p + cc) + j), (uint32_t)g)];
#pragma unroll
            for (int u = 0; u < FT; ++u)
              d[cc][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, hb[s][u], s == 0 ? b0 : d[cc][u], 0, 0, 0);
          }
        }
#pragma unroll
        for (int u = 0; u < FT; ++u) nb[p][u] = next_b(d[0][u], d[1][u]);
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int u = 0; u < FT; ++u) hb[p][u] = nb[p][u];
    }
    // layer 5 (3 outputs): tile u's outputs are placed in lane group u by
    // reading W5's rows shifted by 4u (rows 3..15 of W5 are zero), so every lane
    // ends with the 3 outputs of one row and evaluates 3 exponentials
    const float* b5 = A.b[MLP_HIDDEN_LAYERS];
    const f32x4 y0 = f32x4{b5[0], b5[1], b5[2], 0.0f};
    f32x4 y[FT];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int u = 0; u < FT; ++u) {
        const bf16x8 a =
            *(const bf16x8*)&flds[FW5 + fchunk(16u, (uint32_t)s, (uint32_t)((j - 4 * u) & 15), (uint32_t)g)];
        y[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, hb[s][u], s == 0 ? y0 : y[u], 0, 0, 0);
      }
    const f32x4 yg = g == 0 ? y[0] : (g == 1 ? y[1] : (g == 2 ? y[2] : y[3]));
    const uint32_t row = row0 + 16u * (uint32_t)g + (uint32_t)j;
    if (row < A.rows) {
      const uint32_t orec = A.list ? A.list[row] : row;
#pragma unroll
      for (int e = 0; e < MLP_OUT; ++e) {
        const float v = exp_(yg[e]) - 1.0f;
        A.rgb[(size_t)orec * MLP_OUT + e] = v > 0.0f ? v : 0.0f;
      }
    }
  }
}

size_t mlp_fused_lds_bytes() { return FLDS; }

// Inference for n rows (or the n records of `list`) through the single fused
// kernel (no activation buffer); `volume`: every row is a volume record.
hipError_t mlp_infer_fused(const MlpDev& w, const float* in, uint32_t n, float* rgb, hipStream_t st,
                           const uint32_t* list, bool volume) {
  if (n == 0) return hipSuccess;
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  if (cus <= 0) cus = 1;
  FusedArgs a;
  a.in = in;
  a.rows = n;
  a.list = list;
  for (int l = 0; l < MLP_LAYERS; ++l) {
    a.w[l] = w.wb[l];
    a.b[l] = w.b[l];
  }
  a.rgb = rgb;
  const uint32_t steps = (n + FROWS - 1) / FROWS;
  const uint32_t grid = std::min<uint32_t>((uint32_t)cus, (steps + IWF / 64 - 1) / (IWF / 64));
  if (volume) hipLaunchKernelGGL(k_mlp_fused<true>, dim3(grid), dim3(IWF), FLDS, st, a);
  else hipLaunchKernelGGL(k_mlp_fused<false>, dim3(grid), dim3(IWF), FLDS, st, a);
  return hipGetLastError();
}

size_t mlp_infer_lds_bytes_x() { return (size_t)((3 * MLP_HID + MLP_OUTPAD) * WHS + (IW25 / 64) * 16 * WHS) * 2; }

size_t mlp_infer_h1_rows() { return 1u << 20; }
size_t mlp_infer_lds_bytes() {
  return (size_t)((3 * MLP_HID + MLP_OUTPAD) * WHS + (IW / 64) * ROWS_PER_WAVE * WHS) * 2;
}

// Inference for n rows: h1 must hold mlp_infer_h1_rows() x 128 BF16.
hipError_t mlp_infer(const MlpDev& w, const float* in, uint32_t n, uint16_t* h1, float* rgb, hipStream_t st,
                     int k32, uint32_t chunk_rows) {
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  if (cus <= 0) cus = 1;
  const size_t lds25 = mlp_infer_lds_bytes();
  const uint32_t chunk = std::max<uint32_t>(1024u, std::min<uint32_t>(chunk_rows, (uint32_t)mlp_infer_h1_rows()));
  for (uint32_t o = 0; o < n; o += chunk) {
    const uint32_t rows = std::min(chunk, n - o);
    const uint32_t tiles = (rows + ROWS_PER_WAVE - 1) / ROWS_PER_WAVE;
    const uint32_t want = (tiles + IW / 64 - 1) / (IW / 64);
    uint32_t g1 = std::min<uint32_t>((uint32_t)cus * 2u, want), g2 = std::min<uint32_t>((uint32_t)cus, want);
    const uint32_t want25 = (tiles + IW25 / 64 - 1) / (IW25 / 64);
    const uint32_t g25 = std::min<uint32_t>((uint32_t)cus, want25);
    const uint32_t g1x = std::min<uint32_t>((uint32_t)cus, (tiles + IW1X / 64 - 1) / (IW1X / 64));
    if (k32) hipLaunchKernelGGL(k_mlp_l1x, dim3(g1x), dim3(IW1X), 0, st, in + (size_t)o * MLP_IN, rows, w.wb[0], w.b[0], h1);
    else hipLaunchKernelGGL(k_mlp_l1, dim3(g1), dim3(IW), 0, st, in + (size_t)o * MLP_IN, rows, w.wb[0], w.b[0], h1);
    L25Args a;
    a.h1 = h1;
    a.rows = rows;
    for (int l = 0; l < 4; ++l) {
      a.w[l] = w.wb[l + 1];
      a.b[l] = w.b[l + 1];
    }
    a.rgb = rgb + (size_t)o * MLP_OUT;
    if (k32) hipLaunchKernelGGL(k_mlp_l25x, dim3(g25), dim3(IW25), mlp_infer_lds_bytes_x(), st, a);
    else hipLaunchKernelGGL(k_mlp_l25, dim3(g2), dim3(IW), lds25, st, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace pm
