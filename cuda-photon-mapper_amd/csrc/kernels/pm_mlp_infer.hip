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
  __shared__ uint16_t W[MLP_HID * W1S];
  for (int e = threadIdx.x; e < MLP_HID * MLP_KPAD; e += IW) {
    int n = e / MLP_KPAD, k = e % MLP_KPAD;
    W[n * W1S + k] = w1[e];
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
    for (int e = threadIdx.x; e < nout * MLP_HID; e += IW) {
      int n = e / MLP_HID, k = e % MLP_HID;
      W[(l * MLP_HID + n) * WHS + k] = A.w[l][e];
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

size_t mlp_infer_h1_rows() { return 1u << 18; }
size_t mlp_infer_lds_bytes() {
  return (size_t)((3 * MLP_HID + MLP_OUTPAD) * WHS + (IW / 64) * ROWS_PER_WAVE * WHS) * 2;
}

// Inference for n rows: h1 must hold mlp_infer_h1_rows() x 128 BF16.
hipError_t mlp_infer(const MlpDev& w, const float* in, uint32_t n, uint16_t* h1, float* rgb, hipStream_t st) {
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  if (cus <= 0) cus = 1;
  const size_t lds25 = mlp_infer_lds_bytes();
  const uint32_t chunk = (uint32_t)mlp_infer_h1_rows();
  for (uint32_t o = 0; o < n; o += chunk) {
    const uint32_t rows = std::min(chunk, n - o);
    const uint32_t tiles = (rows + ROWS_PER_WAVE - 1) / ROWS_PER_WAVE;
    const uint32_t want = (tiles + IW / 64 - 1) / (IW / 64);
    uint32_t g1 = std::min<uint32_t>((uint32_t)cus * 2u, want), g2 = std::min<uint32_t>((uint32_t)cus, want);
    hipLaunchKernelGGL(k_mlp_l1, dim3(g1), dim3(IW), 0, st, in + (size_t)o * MLP_IN, rows, w.wb[0], w.b[0], h1);
    L25Args a;
    a.h1 = h1;
    a.rows = rows;
    for (int l = 0; l < 4; ++l) {
      a.w[l] = w.wb[l + 1];
      a.b[l] = w.b[l + 1];
    }
    a.rgb = rgb + (size_t)o * MLP_OUT;
    hipLaunchKernelGGL(k_mlp_l25, dim3(g2), dim3(IW), lds25, st, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace pm
