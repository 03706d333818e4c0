// pm_mlp.hip -- learned radiance cache: frequency encoding, the general BF16
// MFMA GEMM (FP32 accumulation) of the unfused inference path, and Adam, for
// gfx950.  The training pass is in pm_mlp_train.hip, the fused inference in
// pm_mlp_infer.hip.
//
// No reference counterpart: the reference has no learned component (its
// radiance comes only from the voxel-grid sums, photonMappingKernel.cu:
// 314-389, 831-870).  The cache predicts the same quantity as the kNN gather
// (surface outgoing radiance or volume in-scatter) from
//   input  = position (3) + direction (3) + normal or zero (3)
//   encode = sin/cos(2^b pi v), b = 0..11  -> 216 features (padded to 224)
//   MLP    = 4 x (128, ReLU) -> 3, output = log(1 + radiance)
//
// GEMM: C[M x N] = A[M x K] * B[K x N], A/B in BF16 (either may be read
// transposed), FP32 accumulation on v_mfma_f32_16x16x16_bf16.  A 256-thread
// workgroup computes a 64x64 tile of C; each wave a 32x32 quarter (2x2 MFMA
// tiles); K advances 32 at a time through LDS.  With grid.z > 1 it splits K
// over slices that write FP32 partials (slice z at Cf + z * slice_stride).
#include "pm_kernels.h"
#include "pm_mlp.h"
#include "pm_wave.h"

namespace pm {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) { return f32_to_bf16(f); }
__device__ __forceinline__ float bf2f(uint16_t h) { return bf16_to_f32(h); }

// ---------------------------------------------------------------- encoding
// out[r][i*24 + 2b + {0,1}] = {sin, cos}(2 pi * 2^(b-1) * in[r][i]); padded to MLP_KPAD.
__global__ void k_encode(const float* __restrict__ in, uint32_t n, uint16_t* __restrict__ out) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t r = t / MLP_IN, i = t % MLP_IN;
  if (r >= n) return;
  float v = in[(size_t)r * MLP_IN + i];
  uint16_t* o = out + (size_t)r * MLP_KPAD + i * 2 * MLP_BANDS;
  float scale = 0.5f;
  for (int b = 0; b < MLP_BANDS; ++b) {
    float s, c;
    sincos_turns(v * scale, &s, &c);
    o[2 * b] = f2bf(s);
    o[2 * b + 1] = f2bf(c);
    scale *= 2.0f;
  }
  if (i == MLP_IN - 1)
    for (int k = MLP_FEAT; k < MLP_KPAD; ++k) out[(size_t)r * MLP_KPAD + k] = 0;
}

// ---------------------------------------------------------------- GEMM
static constexpr int GT = 64;   // tile M and N
static constexpr int GK = 32;   // K per LDS stage
static constexpr int GPAD = 8;  // LDS row padding (bf16 elements)

struct GemmArgs {
  int M, N, K;
  const uint16_t* A;
  int lda, transA;  // A(m,k) = transA ? A[k*lda + m] : A[m*lda + k]
  const uint16_t* B;
  int ldb, transB;  // B(k,n) = transB ? B[n*ldb + k] : B[k*ldb + n]
  float* Cf;        // FP32 output (or split-K partials, slice z at Cf + z * slice_stride)
  uint16_t* Cb;     // BF16 output
  int ldc;
  size_t slice_stride;
  const float* bias;       // per column n
  int relu;                // max(0, .)
  const uint16_t* mask;    // multiply by (mask(m,n) > 0), mask row-major with ldc
  int k_per_slice;         // split-K slice length (multiple of GK)
};

__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
  __shared__ uint16_t As[GT][GK + GPAD];  // [m][k]
  __shared__ uint16_t Bs[GK][GT + GPAD];  // [k][n]
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wave = tid >> 6;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int k_begin = blockIdx.z * g.k_per_slice;
  const int k_end = min(g.K, k_begin + g.k_per_slice);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = k_begin; k0 < k_end; k0 += GK) {
    // stage A (64 x 32) and B (32 x 64): 2048 elements each = one run of 8
    // elements along the operand's contiguous dimension per thread, loaded
    // with one 16-byte load when the run is in range and aligned (the
    // training shapes), element by element otherwise
    {
      const int run = tid;  // 256 runs of 8
      int mm, kk;           // first element of the run
      if (g.transA) { mm = (run % (GT / 8)) * 8; kk = run / (GT / 8); } else { kk = (run % (GK / 8)) * 8; mm = run / (GK / 8); }
      const int gm = m0 + mm, gk = k0 + kk;
      uint16_t v[8];
      const size_t base = g.transA ? (size_t)gk * g.lda + gm : (size_t)gm * g.lda + gk;
      const bool full = g.transA ? (gk < k_end && gm + 7 < g.M) : (gm < g.M && gk + 7 < k_end);
      if (full && ((uintptr_t)(g.A + base) & 15u) == 0) {
        uint4 q = *(const uint4*)&g.A[base];
        v[0] = (uint16_t)(q.x & 0xffffu); v[1] = (uint16_t)(q.x >> 16);
        v[2] = (uint16_t)(q.y & 0xffffu); v[3] = (uint16_t)(q.y >> 16);
        v[4] = (uint16_t)(q.z & 0xffffu); v[5] = (uint16_t)(q.z >> 16);
        v[6] = (uint16_t)(q.w & 0xffffu); v[7] = (uint16_t)(q.w >> 16);
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int em = g.transA ? gm + j : gm, ek = g.transA ? gk : gk + j;
          v[j] = (em < g.M && ek < k_end) ? g.A[g.transA ? (size_t)ek * g.lda + em : (size_t)em * g.lda + ek] : 0;
        }
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        if (g.transA) As[mm + j][kk] = v[j];
        else As[mm][kk + j] = v[j];
      }
    }
    {
      const int run = tid;
      int kk, nn;
      if (g.transB) { kk = (run % (GK / 8)) * 8; nn = run / (GK / 8); } else { nn = (run % (GT / 8)) * 8; kk = run / (GT / 8); }
      const int gk = k0 + kk, gn = n0 + nn;
      uint16_t v[8];
      const size_t base = g.transB ? (size_t)gn * g.ldb + gk : (size_t)gk * g.ldb + gn;
      const bool full = g.transB ? (gn < g.N && gk + 7 < k_end) : (gk < k_end && gn + 7 < g.N);
      if (full && ((uintptr_t)(g.B + base) & 15u) == 0) {
        uint4 q = *(const uint4*)&g.B[base];
        v[0] = (uint16_t)(q.x & 0xffffu); v[1] = (uint16_t)(q.x >> 16);
        v[2] = (uint16_t)(q.y & 0xffffu); v[3] = (uint16_t)(q.y >> 16);
        v[4] = (uint16_t)(q.z & 0xffffu); v[5] = (uint16_t)(q.z >> 16);
        v[6] = (uint16_t)(q.w & 0xffffu); v[7] = (uint16_t)(q.w >> 16);
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int ek = g.transB ? gk + j : gk, en = g.transB ? gn : gn + j;
          v[j] = (en < g.N && ek < k_end) ? g.B[g.transB ? (size_t)en * g.ldb + ek : (size_t)ek * g.ldb + en] : 0;
        }
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        if (g.transB) Bs[kk + j][nn] = v[j];
        else Bs[kk][nn + j] = v[j];
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < GK; ks += 16) {
      bf16x4 af[2], bfr[2];
      const int kq = ks + 4 * (lane >> 4);
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int row = wm + t * 16 + (lane & 15);
        af[t] = bf16x4{(short)As[row][kq], (short)As[row][kq + 1], (short)As[row][kq + 2], (short)As[row][kq + 3]};
        const int col = wn + t * 16 + (lane & 15);
        bfr[t] = bf16x4{(short)Bs[kq][col], (short)Bs[kq + 1][col], (short)Bs[kq + 2][col], (short)Bs[kq + 3][col]};
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[a], bfr[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
  // epilogue: lane holds D[4*(lane/16) + e][lane%16] of each 16x16 tile
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int gm = m0 + wm + a * 16 + 4 * (lane >> 4) + e;
        const int gn = n0 + wn + b * 16 + (lane & 15);
        if (gm >= g.M || gn >= g.N) continue;
        float v = acc[a][b][e];
        if (g.bias) v += g.bias[gn];
        if (g.relu) v = v > 0.0f ? v : 0.0f;
        if (g.mask && !(bf2f(g.mask[(size_t)gm * g.ldc + gn]) > 0.0f)) v = 0.0f;
        if (g.Cf) g.Cf[blockIdx.z * g.slice_stride + (size_t)gm * g.ldc + gn] = v;
        if (g.Cb) g.Cb[(size_t)gm * g.ldc + gn] = f2bf(v);
      }
}

// Adam on FP32 master parameters, refreshing the BF16 copy.
// Adam over every weight and bias tensor in one launch: tensor s takes the
// blocks [first[s], first[s + 1]).
struct AdamSegs {
  float* p[2 * MLP_LAYERS];
  float* m[2 * MLP_LAYERS];
  float* v[2 * MLP_LAYERS];
  const float* g[2 * MLP_LAYERS];
  uint16_t* pb[2 * MLP_LAYERS];
  uint16_t* pbt[2 * MLP_LAYERS];  // transposed BF16 copy [kin x nout] of a weight tensor, or null
  uint32_t kin[2 * MLP_LAYERS], nout[2 * MLP_LAYERS];
  uint32_t n[2 * MLP_LAYERS];
  const float* gp[2 * MLP_LAYERS];  // split partials (slice z at gp + z * n), when slices > 0
  int slices;
  float* pe[2 * MLP_LAYERS];        // moving average of the tensor (null: none)
  uint16_t* pbe[2 * MLP_LAYERS];    // its BF16 copy (weights only)
  float ema_decay;
  uint32_t first[2 * MLP_LAYERS + 1];
};

__global__ void k_adam_all(AdamSegs S, float lr, float b1, float b2, float eps, float c1, float c2) {
  int s = 0;
  while (s + 1 < 2 * MLP_LAYERS && blockIdx.x >= S.first[s + 1]) ++s;
  const size_t i = (size_t)(blockIdx.x - S.first[s]) * blockDim.x + threadIdx.x;
  if (i >= S.n[s]) return;
  float* p = S.p[s];
  float* m = S.m[s];
  float* v = S.v[s];
  float gr;
  if (S.slices > 0) {  // the same fixed-order sum over slices as k_grad_sum
    gr = 0.0f;
    for (int z = 0; z < S.slices; ++z) gr += S.gp[s][(size_t)z * S.n[s] + i];
  } else {
    gr = S.g[s][i];
  }
  float mi = b1 * m[i] + (1.0f - b1) * gr;
  float vi = b2 * v[i] + (1.0f - b2) * gr * gr;
  m[i] = mi;
  v[i] = vi;
  float mh = mi / c1, vh = vi / c2;
  float pi = p[i] - lr * mh / (sqrt_(vh) + eps);
  p[i] = pi;
  if (S.pb[s]) S.pb[s][i] = f2bf(pi);
  if (S.pbt[s]) S.pbt[s][(i % S.kin[s]) * S.nout[s] + i / S.kin[s]] = f2bf(pi);
  if (S.pe[s]) {
    const float e = S.ema_decay * S.pe[s][i] + (1.0f - S.ema_decay) * pi;
    S.pe[s][i] = e;
    if (S.pbe[s]) S.pbe[s][i] = f2bf(e);
  }
}

// BF16 copy of an [nout x kin] weight matrix, and its transpose when yt is set.
__global__ void k_to_bf16(const float* __restrict__ x, size_t n, uint16_t* __restrict__ y, uint16_t* __restrict__ yt,
                          uint32_t kin, uint32_t nout) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint16_t h = f2bf(x[i]);
  y[i] = h;
  if (yt) yt[(i % kin) * nout + i / kin] = h;
}

// radiance = max(0, exp(y) - 1) for the first 3 output columns.
__global__ void k_decode_out(const float* __restrict__ y, size_t rows, float* __restrict__ rgb) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  for (int c = 0; c < MLP_OUT; ++c) {
    float e = exp_(y[r * MLP_OUTPAD + c]) - 1.0f;
    rgb[r * MLP_OUT + c] = e > 0.0f ? e : 0.0f;
  }
}

// ---------------------------------------------------------------- host side
static hipError_t gemm(const GemmArgs& a, int slices, hipStream_t st) {
  dim3 grid((unsigned)((a.N + GT - 1) / GT), (unsigned)((a.M + GT - 1) / GT), (unsigned)slices);
  hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, st, a);
  return hipGetLastError();
}

static GemmArgs gemm_args(int M, int N, int K) {
  GemmArgs a;
  memset(&a, 0, sizeof(a));
  a.M = M;
  a.N = N;
  a.K = K;
  a.k_per_slice = ((K + GK - 1) / GK) * GK;
  return a;
}

hipError_t mlp_encode(const float* in, uint32_t n, uint16_t* out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  uint64_t t = (uint64_t)n * MLP_IN;
  hipLaunchKernelGGL(k_encode, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, st, in, n, out);
  return hipGetLastError();
}

// Forward for `rows` encoded inputs X [rows x KPAD]; H[l] receives the
// post-ReLU activations (BF16), y the FP32 output [rows x OUTPAD].
hipError_t mlp_forward(const MlpDev& w, const uint16_t* X, uint32_t rows, uint16_t* const* H, float* y,
                       hipStream_t st) {
  hipError_t e;
  const uint16_t* in = X;
  int kin = MLP_KPAD;
  for (int l = 0; l < MLP_HIDDEN_LAYERS; ++l) {
    GemmArgs a = gemm_args((int)rows, MLP_HID, kin);
    a.A = in; a.lda = kin; a.transA = 0;
    a.B = w.wb[l]; a.ldb = kin; a.transB = 1;  // W [out x in] row-major -> B(k, n) = W[n][k]
    a.Cb = H[l]; a.ldc = MLP_HID;
    a.bias = w.b[l];
    a.relu = 1;
    if ((e = gemm(a, 1, st)) != hipSuccess) return e;
    in = H[l];
    kin = MLP_HID;
  }
  GemmArgs a = gemm_args((int)rows, MLP_OUTPAD, MLP_HID);
  a.A = in; a.lda = MLP_HID;
  a.B = w.wb[MLP_HIDDEN_LAYERS]; a.ldb = MLP_HID; a.transB = 1;
  a.Cf = y; a.ldc = MLP_OUTPAD;
  a.bias = w.b[MLP_HIDDEN_LAYERS];
  return gemm(a, 1, st);
}

hipError_t mlp_decode(const float* y, uint32_t rows, float* rgb, hipStream_t st) {
  if (rows == 0) return hipSuccess;
  hipLaunchKernelGGL(k_decode_out, dim3((rows + 255) / 256), dim3(256), 0, st, y, (size_t)rows, rgb);
  return hipGetLastError();
}

hipError_t mlp_adam(MlpDev& w, const MlpTrainBufs& tb, float lr, int step, hipStream_t st, const MlpTrainT* parts) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float c1 = 1.0f - (float)std::pow(0.9, (double)step), c2 = 1.0f - (float)std::pow(0.999, (double)step);
  AdamSegs S;
  uint32_t blocks = 0;
  for (int l = 0; l <= MLP_HIDDEN_LAYERS; ++l) {
    for (int j = 0; j < 2; ++j) {
      const int s = 2 * l + j;
      S.p[s] = j ? w.b[l] : w.w[l];
      S.m[s] = j ? w.mb[l] : w.mw[l];
      S.v[s] = j ? w.vb[l] : w.vw[l];
      S.g[s] = j ? tb.gb[l] : tb.gw[l];
      S.pb[s] = j ? nullptr : w.wb[l];
      S.pbt[s] = j ? nullptr : w.wbt[l];
      S.kin[s] = (uint32_t)mlp_kin(l);
      S.nout[s] = (uint32_t)mlp_nout(l);
      S.n[s] = (uint32_t)(j ? mlp_b_count(l) : mlp_w_count(l));
      S.gp[s] = parts ? (j ? parts->part_b[l] : parts->part_w[l]) : nullptr;
      S.pe[s] = w.ema_decay > 0.0f ? (j ? w.be[l] : w.we[l]) : nullptr;
      S.pbe[s] = w.ema_decay > 0.0f && !j ? w.wbe[l] : nullptr;
      S.first[s] = blocks;
      blocks += (S.n[s] + 255u) / 256u;
    }
  }
  S.first[2 * MLP_LAYERS] = blocks;
  S.slices = parts ? parts->slices : 0;
  S.ema_decay = w.ema_decay;
  hipLaunchKernelGGL(k_adam_all, dim3(blocks), dim3(256), 0, st, S, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

MlpDev mlp_inference_net(const MlpDev& w) {
  MlpDev n = w;
  if (w.ema_decay > 0.0f)
    for (int l = 0; l < MLP_LAYERS; ++l) {
      n.w[l] = w.we[l];
      n.b[l] = w.be[l];
      n.wb[l] = w.wbe[l];
    }
  return n;
}

hipError_t mlp_reset_ema(MlpDev& w, hipStream_t st) {
  for (int l = 0; l < MLP_LAYERS; ++l) {
    hipError_t e = hipMemcpyAsync(w.we[l], w.w[l], 4 * mlp_w_count(l), hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(w.be[l], w.b[l], 4 * mlp_b_count(l), hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(w.wbe[l], w.wb[l], 2 * mlp_w_count(l), hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t mlp_refresh_bf16(MlpDev& w, hipStream_t st) {
  for (int l = 0; l <= MLP_HIDDEN_LAYERS; ++l) {
    size_t nw = mlp_w_count(l);
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, st, w.w[l], nw, w.wb[l],
                       w.wbt[l], (uint32_t)mlp_kin(l), (uint32_t)mlp_nout(l));
  }
  return hipGetLastError();
}

}  // namespace pm
