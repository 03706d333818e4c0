// pm_mlp_train.hip -- the learned cache's training pass (forward with stored
// activations, backward, weight and bias gradients) on gfx950: the forward
// pass and loss gradient in one fused kernel (k_fwd_fused), the backward
// products on one "NT" GEMM kernel whose operands are both contiguous along K.
//
// No reference counterpart (the reference has no learned component).
//
// Every product of a training step has the form C[m][n] = sum_k A[m][k] B[n][k]
// once each operand is kept in the layout that makes K its contiguous axis:
//   forward      H_l [rows x out] = X_{l} [rows x in] . W_l [out x in]^T
//   delta        D_{l-1} [rows x in] = D_l [rows x out] . (W_l^T) [in x out]^T
//   weight grad  dW_l [out x in] = (D_l^T) [out x rows] . (X_l^T) [in x rows]^T
// so the forward and delta products also write their output transposed (the
// weight gradients' operands), the encoder writes X and X^T, and the BF16
// weights are kept in both orders (W and W^T, refreshed by Adam).  A
// 256-thread workgroup computes a 64 x 64 tile of C; each wave a 32 x 32
// quarter on v_mfma_f32_16x16x32_bf16.  K is staged through LDS 128 at a time
// with 16-byte loads and 16-byte LDS fragment reads.  The weight gradients
// split the batch (K) over grid.z slices that write FP32 partial tiles, and
// the bias gradients (row sums of D^T) likewise; the slices of every layer are
// added in slice order (deterministic), by one kernel or by Adam itself.
#include "pm_kernels.h"
#include "pm_mlp.h"
#include "pm_wave.h"

namespace pm {

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int NT_TILE = 64;  // tile M and N
constexpr int NT_KC = 128;   // K per LDS stage
constexpr int NT_LDS = NT_KC + 8;  // LDS row length (bf16), padded

__device__ __forceinline__ float bf_to_f(uint16_t h) { return bf16_to_f32(h); }

struct NtArgs {
  int M, N, K;
  const uint16_t* A;  // A[m * lda + k]
  int lda;
  const uint16_t* B;  // B[n * ldb + k]
  int ldb;
  // epilogue of an unsplit product
  const float* bias;       // + bias[n]
  int relu;                // max(0, .)
  const uint16_t* mask;    // * (mask[m * ldm + n] > 0)
  int ldm;
  uint16_t* Cb;            // BF16 C[m * ldc + n]
  int ldc;
  uint16_t* Ct;            // BF16 C^T[n * ldt + m] (ldt >= M rounded up to 64)
  int ldt;
  float* Cf;               // FP32 C[m * ldf + n]
  int ldf;
  // split product (gridDim.z slices of kps): partial tiles part[z][m][n],
  // and when bpart is set the row sums of A over the slice, bpart[z][m]
  int kps;
  float* part;
  float* bpart;
};

// Stages A rows [m0, m0+64) and B rows [n0, n0+64), k in [k0, k0 + kc) (kc a
// multiple of 32), into LDS; elements outside M, N or k_end are zero.
__device__ __forceinline__ void nt_stage(const NtArgs& g, uint16_t (*As)[NT_LDS], uint16_t (*Bs)[NT_LDS], int m0,
                                         int n0, int k0, int kc, int k_end) {
  const int vec_per_row = kc / 8;
  for (int v = (int)threadIdx.x; v < 2 * NT_TILE * vec_per_row; v += 256) {
    const bool isB = v >= NT_TILE * vec_per_row;
    const int w = isB ? v - NT_TILE * vec_per_row : v;
    const int r = w / vec_per_row, kk = (w % vec_per_row) * 8;
    const int gr = (isB ? n0 : m0) + r, gk = k0 + kk;
    const int rlim = isB ? g.N : g.M;
    const uint16_t* src = isB ? g.B + (size_t)gr * g.ldb + gk : g.A + (size_t)gr * g.lda + gk;
    uint4 q = make_uint4(0u, 0u, 0u, 0u);
    if (gr < rlim) {
      if (gk + 8 <= k_end) {
        q = *(const uint4*)src;
      } else if (gk < k_end) {
        uint16_t e[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) e[j] = gk + j < k_end ? src[j] : (uint16_t)0;
        q.x = (uint32_t)e[0] | ((uint32_t)e[1] << 16);
        q.y = (uint32_t)e[2] | ((uint32_t)e[3] << 16);
        q.z = (uint32_t)e[4] | ((uint32_t)e[5] << 16);
        q.w = (uint32_t)e[6] | ((uint32_t)e[7] << 16);
      }
    }
    uint16_t* dst = isB ? &Bs[r][kk] : &As[r][kk];
    *(uint4*)dst = q;
  }
}

// Tile (bx, by) of slice bz (of nz) of the product g.
__device__ __forceinline__ void nt_tile(const NtArgs& g, int bx, int by, int bz, int nz, uint16_t (*As)[NT_LDS],
                                        uint16_t (*Bs)[NT_LDS]) {
  const int lane = lane_id();
  const int wave = (int)(threadIdx.x >> 6);
  const int m0 = by * NT_TILE, n0 = bx * NT_TILE;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const bool split = nz > 1 || g.part;
  const int k_begin = split ? bz * g.kps : 0;
  const int k_end = split ? min(g.K, k_begin + g.kps) : g.K;
  const int r16 = lane & 15, q = lane >> 4;
  f32x4_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
  float rowsum = 0.0f;  // bias gradient: row threadIdx.x of A over this slice (n-tile 0)
  const bool want_rowsum = g.bpart && bx == 0 && threadIdx.x < NT_TILE;
  // the epilogue's bias and mask values, loaded before the K loop so that
  // their latency overlaps the operand loads
  float bias_v[2] = {0.0f, 0.0f};
  uint16_t mask_v[2][2][4];
  if (!split) {
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int n = n0 + wn + 16 * b + r16;
      if (g.bias && n < g.N) bias_v[b] = g.bias[n];
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int m = m0 + wm + 16 * a + 4 * q + e, n = n0 + wn + 16 * b + r16;
          // 0x3f80 is 1.0 in BF16: kept when there is no mask
          mask_v[a][b][e] = g.mask && m < g.M && n < g.N ? g.mask[(size_t)m * g.ldm + n] : (uint16_t)0x3f80;
        }
  }
  for (int k0 = k_begin; k0 < k_end; k0 += NT_KC) {
    const int kc = min(NT_KC, (k_end - k0 + 31) / 32 * 32);
    nt_stage(g, As, Bs, m0, n0, k0, kc, k_end);
    __syncthreads();
    if (want_rowsum)
      for (int k = 0; k < kc; ++k) rowsum += bf_to_f(As[threadIdx.x][k]);
    for (int ks = 0; ks < kc; ks += 32) {
      bf16x8_t af[2], bfr[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const uint4 va = *(const uint4*)&As[wm + t * 16 + r16][ks + 8 * q];
        af[t] = *(const bf16x8_t*)&va;
        const uint4 vb = *(const uint4*)&Bs[wn + t * 16 + r16][ks + 8 * q];
        bfr[t] = *(const bf16x8_t*)&vb;
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
  // lane holds C[wm + 16a + 4q + e][wn + 16b + r16]
  if (split) {
    float* P = g.part + (size_t)bz * g.M * g.N;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int m = m0 + wm + 16 * a + 4 * q + e, n = n0 + wn + 16 * b + r16;
          if (m < g.M && n < g.N) P[(size_t)m * g.N + n] = acc[a][b][e];
        }
    if (want_rowsum && m0 + (int)threadIdx.x < g.M)
      g.bpart[(size_t)bz * g.M + m0 + threadIdx.x] = rowsum;
    return;
  }
  // unsplit: bias, ReLU, mask, then the BF16 / FP32 outputs; the transposed
  // BF16 output goes through LDS so its rows are written 16 bytes at a time
  uint16_t (*T)[NT_TILE + 8] = (uint16_t (*)[NT_TILE + 8])&As[0][0];  // [n][m], 9.2 KB of As
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int ml = wm + 16 * a + 4 * q + e, nl = wn + 16 * b + r16;
        const int m = m0 + ml, n = n0 + nl;
        float v = acc[a][b][e];
        if (n < g.N) {
          if (g.bias) v += bias_v[b];
          if (g.relu) v = v > 0.0f ? v : 0.0f;
          if (!(bf_to_f(mask_v[a][b][e]) > 0.0f)) v = 0.0f;
        }
        const uint16_t h = f32_to_bf16(v);
        if (m < g.M && n < g.N) {
          if (g.Cb) g.Cb[(size_t)m * g.ldc + n] = h;
          if (g.Cf) g.Cf[(size_t)m * g.ldf + n] = v;
        }
        if (g.Ct) T[nl][ml] = h;
      }
  if (g.Ct) {
    __syncthreads();
    // 64 rows (n) x 8 vectors of 8 m; columns m >= M land in the padding of C^T
    for (int v = (int)threadIdx.x; v < NT_TILE * 8; v += 256) {
      const int nl = v >> 3, j = (v & 7) * 8;
      if (n0 + nl < g.N) *(uint4*)&g.Ct[(size_t)(n0 + nl) * g.ldt + m0 + j] = *(const uint4*)&T[nl][j];
    }
  }
}

__global__ void __launch_bounds__(256) k_gemm_nt(NtArgs g) {
  __shared__ __align__(16) uint16_t As[NT_TILE][NT_LDS];
  __shared__ __align__(16) uint16_t Bs[NT_TILE][NT_LDS];
  nt_tile(g, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.z, As, Bs);
}

// Independent products in one launch (the five weight gradients of a batch):
// workgroups [first[s], first[s + 1]) take product s, n-tiles fastest, then
// m-tiles, then slices.
struct NtMulti {
  NtArgs a[MLP_LAYERS];
  int nx[MLP_LAYERS], ny[MLP_LAYERS], nz[MLP_LAYERS];
  int first[MLP_LAYERS + 1];
};

template <int S>
__device__ __forceinline__ void nt_multi_tile(const NtMulti& p, int c, uint16_t (*As)[NT_LDS], uint16_t (*Bs)[NT_LDS]) {
  nt_tile(p.a[S], c % p.nx[S], (c / p.nx[S]) % p.ny[S], c / (p.nx[S] * p.ny[S]), p.nz[S], As, Bs);
}

__global__ void __launch_bounds__(256) k_gemm_nt_multi(NtMulti p) {
  __shared__ __align__(16) uint16_t As[NT_TILE][NT_LDS];
  __shared__ __align__(16) uint16_t Bs[NT_TILE][NT_LDS];
  const int b = (int)blockIdx.x;
  int s = 0;
  while (s + 1 < MLP_LAYERS && b >= p.first[s + 1]) ++s;
  const int c = b - p.first[s];
  static_assert(MLP_LAYERS == 5, "one case per layer");
  switch (s) {
    case 0: nt_multi_tile<0>(p, c, As, Bs); break;
    case 1: nt_multi_tile<1>(p, c, As, Bs); break;
    case 2: nt_multi_tile<2>(p, c, As, Bs); break;
    case 3: nt_multi_tile<3>(p, c, As, Bs); break;
    default: nt_multi_tile<4>(p, c, As, Bs); break;
  }
}

hipError_t gemm_nt(const NtArgs& a, int slices, hipStream_t st) {
  dim3 grid((unsigned)((a.N + NT_TILE - 1) / NT_TILE), (unsigned)((a.M + NT_TILE - 1) / NT_TILE), (unsigned)slices);
  hipLaunchKernelGGL(k_gemm_nt, grid, dim3(256), 0, st, a);
  return hipGetLastError();
}

// The products g[0..MLP_LAYERS) with slices[s] K slices each, one launch.
hipError_t gemm_nt_multi(const NtArgs* g, const int* slices, hipStream_t st) {
  NtMulti p;
  p.first[0] = 0;
  for (int i = 0; i < MLP_LAYERS; ++i) {
    p.a[i] = g[i];
    p.nx[i] = (g[i].N + NT_TILE - 1) / NT_TILE;
    p.ny[i] = (g[i].M + NT_TILE - 1) / NT_TILE;
    p.nz[i] = slices[i];
    p.first[i + 1] = p.first[i] + p.nx[i] * p.ny[i] * p.nz[i];
  }
  hipLaunchKernelGGL(k_gemm_nt_multi, dim3((unsigned)p.first[MLP_LAYERS]), dim3(256), 0, st, p);
  return hipGetLastError();
}

NtArgs nt_args(int M, int N, int K) {
  NtArgs a;
  memset(&a, 0, sizeof(a));
  a.M = M;
  a.N = N;
  a.K = K;
  return a;
}

// ---------------------------------------------------------------- fused forward
// The forward pass of a training batch in one launch: a workgroup takes 64
// rows through the five layers with the activations in LDS, writes each
// hidden layer's output row-major (H_l) and transposed (H_l^T), then the
// output layer's y and the loss gradient: with t = log(1 + max(target, 0)),
// d = y - t per channel, dy = 2 d / (3 rows) (BF16, also transposed; zero in
// the padding columns) and row_err = sum of d^2.  When esum is set, each
// row's row_err (capped at 1e6) is added in 2^-24 fixed point to the error
// tracker's cell of its input, and the cell's count incremented (integer
// atomics: the bins do not depend on the order).  Every
// product is k_gemm_nt's (the same MFMA operands, K in ascending steps of 32
// from a zero accumulator, then bias, ReLU and BF16 rounding), so the outputs
// equal those of one launch per layer bit for bit.  A wave computes the
// tile's 64 rows x 32 of a hidden layer's 128 outputs, or 16 rows of the 16
// output-layer columns.
constexpr int FF_XS = MLP_KPAD + 8;  // LDS row length (bf16) of X and W1
constexpr int FF_HS = MLP_HID + 8;   // of the hidden activations and W2..W5
constexpr size_t FF_LDS = 2u * (64u * FF_XS + 2u * 64u * FF_HS + 128u * FF_XS) + 4u * 64u * MLP_OUTPAD;

struct FwdArgs {
  uint16_t* XT;       // [KPAD x ld] encoded inputs, transposed (written here)
  uint32_t rows;
  const uint16_t* w[MLP_LAYERS];
  const float* b[MLP_LAYERS];
  uint16_t* H[MLP_HIDDEN_LAYERS];   // [rows x HID]
  uint16_t* HT[MLP_HIDDEN_LAYERS];  // [HID x ld]
  uint32_t ld;
  float* y;                         // [rows x OUTPAD]
  const float* target;              // [rows x OUT]
  uint16_t* dy;                     // [rows x OUTPAD]
  uint16_t* dyT;                    // [OUTPAD x ld]
  float* row_err;                   // [rows]
  const float* in;                  // [rows x IN] query rows
  unsigned long long* esum;         // error tracker (null: none)
  unsigned int* ecnt;
};

__global__ void __launch_bounds__(256) k_fwd_fused(FwdArgs A) {
  extern __shared__ __align__(16) unsigned char fl[];
  uint16_t* Xs = (uint16_t*)fl;                  // [64][FF_XS]
  uint16_t* Hs[2] = {Xs + 64 * FF_XS, Xs + 64 * FF_XS + 64 * FF_HS};  // [64][FF_HS] each
  uint16_t* Ws = Xs + 64 * FF_XS + 2 * 64 * FF_HS;  // [128][FF_XS] or [nout][FF_HS]
  float* ys = (float*)(Ws + 128 * FF_XS);        // [64][OUTPAD]
  const int tid = (int)threadIdx.x, lane = lane_id(), wave = tid >> 6, r16 = lane & 15, q = lane >> 4;
  const uint32_t m0 = blockIdx.x * 64u;
  const uint32_t rows = A.rows;
  // the tile's encoded inputs: {sin, cos}(2 pi 2^(b-1) v) of each input v
  // at column 24 i + 2b (+1), padding columns and rows past the batch zero
  for (int v = tid; v < 64 * MLP_IN; v += 256) {
    const int r = v / MLP_IN, i = v % MLP_IN;
    const bool ok = m0 + (uint32_t)r < rows;
    const float x = ok ? A.in[(size_t)(m0 + r) * MLP_IN + i] : 0.0f;
    uint16_t* o = &Xs[r * FF_XS + i * 2 * MLP_BANDS];
    float scale = 0.5f;
    for (int b = 0; b < MLP_BANDS; ++b) {
      float sn, cs;
      sincos_turns(x * scale, &sn, &cs);
      o[2 * b] = ok ? f32_to_bf16(sn) : (uint16_t)0;
      o[2 * b + 1] = ok ? f32_to_bf16(cs) : (uint16_t)0;
      scale *= 2.0f;
    }
    if (i == MLP_IN - 1)
      for (int k = MLP_FEAT; k < MLP_KPAD; ++k) Xs[r * FF_XS + k] = 0;
  }
  __syncthreads();
  // X^T: 64 columns of each of its KPAD rows
  for (int v = tid; v < MLP_KPAD * 8; v += 256) {
    const int f = v >> 3, j = (v & 7) * 8;
    uint32_t p[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
      p[t] = (uint32_t)Xs[(j + 2 * t) * FF_XS + f] | ((uint32_t)Xs[(j + 2 * t + 1) * FF_XS + f] << 16);
    *(uint4*)&A.XT[(size_t)f * A.ld + m0 + j] = make_uint4(p[0], p[1], p[2], p[3]);
  }
  const uint16_t* act = Xs;
  int as = FF_XS, kin = MLP_KPAD;
  for (int l = 0; l < MLP_HIDDEN_LAYERS; ++l) {
    const int ws = l == 0 ? FF_XS : FF_HS;
    __syncthreads();  // the previous layer is done with Ws (and its output tile is complete)
    for (int v = tid; v < MLP_HID * (kin / 8); v += 256) {
      const int r = v / (kin / 8), kk = (v % (kin / 8)) * 8;
      *(uint4*)&Ws[r * ws + kk] = *(const uint4*)&A.w[l][(size_t)r * kin + kk];
    }
    __syncthreads();
    f32x4_t acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    for (int ks = 0; ks < kin; ks += 32) {
      bf16x8_t af[4], bfr[2];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const uint4 v = *(const uint4*)&act[(16 * a + r16) * as + ks + 8 * q];
        af[a] = *(const bf16x8_t*)&v;
      }
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const uint4 v = *(const uint4*)&Ws[(32 * wave + 16 * b + r16) * ws + ks + 8 * q];
        bfr[b] = *(const bf16x8_t*)&v;
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
    }
    // lane holds C[16a + 4q + e][32 wave + 16b + r16]
    uint16_t* out = Hs[l & 1];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int n = 32 * wave + 16 * b + r16;
      const float bias = A.b[l][n];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          float v = acc[a][b][e] + bias;
          v = v > 0.0f ? v : 0.0f;
          out[(16 * a + 4 * q + e) * FF_HS + n] = f32_to_bf16(v);
        }
    }
    __syncthreads();
    // H_l rows of the batch, and H_l^T (64 columns of each of its 128 rows)
    for (int v = tid; v < 64 * (MLP_HID / 8); v += 256) {
      const int r = v / (MLP_HID / 8), j = (v % (MLP_HID / 8)) * 8;
      if (m0 + (uint32_t)r < rows) *(uint4*)&A.H[l][(size_t)(m0 + r) * MLP_HID + j] = *(const uint4*)&out[r * FF_HS + j];
    }
    for (int v = tid; v < MLP_HID * 8; v += 256) {
      const int n = v >> 3, j = (v & 7) * 8;
      uint32_t p[4];
#pragma unroll
      for (int t = 0; t < 4; ++t)
        p[t] = (uint32_t)out[(j + 2 * t) * FF_HS + n] | ((uint32_t)out[(j + 2 * t + 1) * FF_HS + n] << 16);
      *(uint4*)&A.HT[l][(size_t)n * A.ld + m0 + j] = make_uint4(p[0], p[1], p[2], p[3]);
    }
    act = out;
    as = FF_HS;
    kin = MLP_HID;
  }
  // output layer: 16 columns (3 used), no ReLU, FP32 y
  __syncthreads();
  for (int v = tid; v < MLP_OUTPAD * (MLP_HID / 8); v += 256) {
    const int r = v / (MLP_HID / 8), kk = (v % (MLP_HID / 8)) * 8;
    *(uint4*)&Ws[r * FF_HS + kk] = *(const uint4*)&A.w[MLP_HIDDEN_LAYERS][(size_t)r * MLP_HID + kk];
  }
  __syncthreads();
  {
    f32x4_t acc = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    for (int ks = 0; ks < MLP_HID; ks += 32) {
      const uint4 va = *(const uint4*)&act[(16 * wave + r16) * FF_HS + ks + 8 * q];
      const uint4 vb = *(const uint4*)&Ws[r16 * FF_HS + ks + 8 * q];
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*(const bf16x8_t*)&va, *(const bf16x8_t*)&vb, acc, 0, 0, 0);
    }
    const float bias = A.b[MLP_HIDDEN_LAYERS][r16];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int m = 16 * wave + 4 * q + e;
      const float v = acc[e] + bias;
      ys[m * MLP_OUTPAD + r16] = v;
      if (m0 + (uint32_t)m < rows) A.y[(size_t)(m0 + m) * MLP_OUTPAD + r16] = v;
    }
  }
  __syncthreads();
  // loss gradient and error bins, one thread per row
  if (tid < 64 && m0 + (uint32_t)tid < rows) {
    const uint32_t r = m0 + (uint32_t)tid;
    float e2 = 0.0f;
    const float inv = 2.0f / (float)((int)rows * MLP_OUT);
    for (int c = 0; c < MLP_OUTPAD; ++c) {
      float gr = 0.0f;
      if (c < MLP_OUT) {
        const float t = log_(1.0f + fmax_(A.target[(size_t)r * MLP_OUT + c], 0.0f));
        const float d = ys[tid * MLP_OUTPAD + c] - t;
        e2 += d * d;
        gr = d * inv;
      }
      const uint16_t h = f32_to_bf16(gr);
      A.dy[(size_t)r * MLP_OUTPAD + c] = h;
      A.dyT[(size_t)c * A.ld + r] = h;
    }
    A.row_err[r] = e2;
    if (A.esum) {
      const uint32_t cell = err_cell(A.in + (size_t)r * MLP_IN);
      const float e = e2 < 1e6f ? e2 : 1e6f;
      atomicAdd(&A.esum[cell], (unsigned long long)(e * 16777216.0f));
      atomicAdd(&A.ecnt[cell], 1u);
    }
  }
}

// ---------------------------------------------------------------- fused deltas
// The deltas of the four hidden layers of a training batch in one launch: a
// workgroup takes 64 rows from dy down to layer 0, D_{l-1} = (D_l . W_l) x
// [H_{l-1} > 0] with D_4 = dy, the deltas in LDS and each written transposed
// (dT[l - 1], the weight-gradient operand).  Every product is k_gemm_nt's
// (dy's 16 columns zero-padded to one K step of 32, the mask applied to the
// FP32 sum before the BF16 rounding), so the outputs equal those of one launch
// per layer bit for bit.  A wave computes the tile's 64 rows x 32 of the 128
// columns.
constexpr size_t FB_LDS = 2u * (2u * 64u * FF_HS + 128u * FF_HS + 64u * FF_HS);

struct BwdArgs {
  const uint16_t* dy;                    // [rows x OUTPAD]
  uint32_t rows;
  const uint16_t* wbt[MLP_LAYERS];       // W_l^T [kin x nout]
  const uint16_t* H[MLP_HIDDEN_LAYERS];  // [rows x HID], the ReLU masks
  uint16_t* dT[MLP_HIDDEN_LAYERS];       // [HID x ld]
  uint32_t ld;
};

__global__ void __launch_bounds__(256) k_bwd_fused(BwdArgs A) {
  extern __shared__ __align__(16) unsigned char fl[];
  uint16_t* const D0 = (uint16_t*)fl;         // [64][FF_HS]: the deltas, ping-pong
  uint16_t* const D1 = D0 + 64 * FF_HS;
  uint16_t* Ws = D0 + 2 * 64 * FF_HS;         // [128][FF_HS]
  uint16_t* Ms = Ws + 128 * FF_HS;                                 // [64][FF_HS]
  const int tid = (int)threadIdx.x, lane = lane_id(), wave = tid >> 6, r16 = lane & 15, q = lane >> 4;
  const uint32_t m0 = blockIdx.x * 64u;
  const uint32_t rows = A.rows;
  // dy rows of the tile: 16 columns, then zeros to 32 (zero past the batch)
  for (int v = tid; v < 64 * 4; v += 256) {
    const int r = v >> 2, j = (v & 3) * 8;
    uint4 val = make_uint4(0u, 0u, 0u, 0u);
    if (j < MLP_OUTPAD && m0 + (uint32_t)r < rows) val = *(const uint4*)&A.dy[(size_t)(m0 + r) * MLP_OUTPAD + j];
    *(uint4*)&D0[r * FF_HS + j] = val;
  }
  const uint16_t* act = D0;
  for (int l = MLP_HIDDEN_LAYERS; l >= 1; --l) {
    const int nout = l == MLP_HIDDEN_LAYERS ? MLP_OUTPAD : MLP_HID;  // K of this product
    const int kpad = (nout + 31) / 32 * 32;
    __syncthreads();  // the previous step is done with Ws and Ms
    // W_l^T [128 x nout], zero-padded to kpad columns
    for (int v = tid; v < MLP_HID * (kpad / 8); v += 256) {
      const int r = v / (kpad / 8), j = (v % (kpad / 8)) * 8;
      uint4 val = make_uint4(0u, 0u, 0u, 0u);
      if (j < nout) val = *(const uint4*)&A.wbt[l][(size_t)r * nout + j];
      *(uint4*)&Ws[r * FF_HS + j] = val;
    }
    // the mask rows H_{l-1}
    for (int v = tid; v < 64 * (MLP_HID / 8); v += 256) {
      const int r = v / (MLP_HID / 8), j = (v % (MLP_HID / 8)) * 8;
      uint4 val = make_uint4(0u, 0u, 0u, 0u);
      if (m0 + (uint32_t)r < rows) val = *(const uint4*)&A.H[l - 1][(size_t)(m0 + r) * MLP_HID + j];
      *(uint4*)&Ms[r * FF_HS + j] = val;
    }
    __syncthreads();
    f32x4_t acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    for (int ks = 0; ks < kpad; ks += 32) {
      bf16x8_t af[4], bfr[2];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const uint4 v = *(const uint4*)&act[(16 * a + r16) * FF_HS + ks + 8 * q];
        af[a] = *(const bf16x8_t*)&v;
      }
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const uint4 v = *(const uint4*)&Ws[(32 * wave + 16 * b + r16) * FF_HS + ks + 8 * q];
        bfr[b] = *(const bf16x8_t*)&v;
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
    }
    // lane holds D_{l-1}[16a + 4q + e][32 wave + 16b + r16] before the mask
    uint16_t* out = ((MLP_HIDDEN_LAYERS - l) & 1) ? D0 : D1;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int n = 32 * wave + 16 * b + r16;
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int m = 16 * a + 4 * q + e;
          float v = acc[a][b][e];
          if (!(bf_to_f(Ms[m * FF_HS + n]) > 0.0f)) v = 0.0f;
          out[m * FF_HS + n] = f32_to_bf16(v);
        }
    }
    __syncthreads();
    for (int v = tid; v < MLP_HID * 8; v += 256) {
      const int n = v >> 3, j = (v & 7) * 8;
      uint32_t p[4];
#pragma unroll
      for (int t = 0; t < 4; ++t)
        p[t] = (uint32_t)out[(j + 2 * t) * FF_HS + n] | ((uint32_t)out[(j + 2 * t + 1) * FF_HS + n] << 16);
      *(uint4*)&A.dT[l - 1][(size_t)n * A.ld + m0 + j] = make_uint4(p[0], p[1], p[2], p[3]);
    }
    act = out;
  }
}

// The gradients of every layer from the split partials: gw[l][i] = sum over
// slices z (in order) of part_l[z][i], gb[l][m] = sum over z of bpart_l[z][m].
struct GradSum {
  const float* part[MLP_LAYERS];
  const float* bpart[MLP_LAYERS];
  float* gw[MLP_LAYERS];
  float* gb[MLP_LAYERS];
  uint32_t nw[MLP_LAYERS], nb[MLP_LAYERS];
  uint32_t first[MLP_LAYERS + 1];  // first thread of layer l (weights then biases)
  int slices;
};

__global__ void k_grad_sum(GradSum S) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= S.first[MLP_LAYERS]) return;
  int l = 0;
  while (l + 1 < MLP_LAYERS && t >= S.first[l + 1]) ++l;
  const uint32_t i = t - S.first[l];
  float s = 0.0f;
  if (i < S.nw[l]) {
    for (int z = 0; z < S.slices; ++z) s += S.part[l][(size_t)z * S.nw[l] + i];
    S.gw[l][i] = s;
  } else {
    const uint32_t m = i - S.nw[l];
    for (int z = 0; z < S.slices; ++z) s += S.bpart[l][(size_t)z * S.nb[l] + m];
    S.gb[l][m] = s;
  }
}

}  // namespace

size_t mlp_train_partial_floats(uint32_t rows) {
  const int slices = mlp_train_slices(rows);
  size_t n = 0;
  for (int l = 0; l < MLP_LAYERS; ++l) n += (mlp_w_count(l) + mlp_b_count(l)) * (size_t)slices;
  return n;
}

int mlp_train_slices(uint32_t rows) { return (int)std::max<uint32_t>(1u, std::min<uint32_t>(32u, (rows + 255u) / 256u)); }

uint32_t mlp_train_ld(uint32_t rows) { return (rows + 63u) / 64u * 64u; }

hipError_t mlp_train_grads(const MlpDev& w, const float* in, uint32_t rows, uint16_t* const* H,
                           float* y, const float* target, MlpTrainBufs& tb, MlpTrainT& tt, bool reduce,
                           hipStream_t st, unsigned long long* esum, unsigned int* ecnt) {
  if (rows == 0) return hipSuccess;
  const uint32_t ld = tt.ld;
  if (ld < mlp_train_ld(rows)) return hipErrorInvalidValue;
  hipError_t e;
  // encoding, forward (X^T, H_l and H_l^T, y), the loss gradient and the
  // error bins: one launch
  {
    FwdArgs f;
    memset(&f, 0, sizeof(f));
    f.XT = tt.XT;
    f.rows = rows;
    for (int l = 0; l < MLP_LAYERS; ++l) {
      f.w[l] = w.wb[l];
      f.b[l] = w.b[l];
    }
    for (int l = 0; l < MLP_HIDDEN_LAYERS; ++l) {
      f.H[l] = H[l];
      f.HT[l] = tt.HT[l];
    }
    f.ld = ld;
    f.y = y;
    f.target = target;
    f.dy = tb.dy;
    f.dyT = tt.dyT;
    f.row_err = tb.row_err;
    f.in = in;
    f.esum = esum;
    f.ecnt = ecnt;
    hipLaunchKernelGGL(k_fwd_fused, dim3((rows + 63) / 64), dim3(256), FF_LDS, st, f);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  // backward: the four hidden layers' deltas in one launch, then the five
  // split weight-gradient products in one launch
  {
    BwdArgs bw;
    memset(&bw, 0, sizeof(bw));
    bw.dy = tb.dy;
    bw.rows = rows;
    for (int l = 0; l < MLP_LAYERS; ++l) bw.wbt[l] = w.wbt[l];
    for (int l = 0; l < MLP_HIDDEN_LAYERS; ++l) {
      bw.H[l] = H[l];
      bw.dT[l] = tt.dT[l];
    }
    bw.ld = ld;
    hipLaunchKernelGGL(k_bwd_fused, dim3((rows + 63) / 64), dim3(256), FB_LDS, st, bw);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  const int slices = mlp_train_slices(rows);
  const int kps = (int)(((rows + (uint32_t)slices - 1u) / (uint32_t)slices + 31u) / 32u * 32u);
  GradSum S;
  memset(&S, 0, sizeof(S));
  S.slices = slices;
  NtArgs wg[MLP_LAYERS];
  int wsl[MLP_LAYERS];
  float* part = tt.part;  // partials of layer 4 first, then 3, ..., 0
  for (int l = MLP_HIDDEN_LAYERS; l >= 0; --l) {
    const int nout = l == MLP_HIDDEN_LAYERS ? MLP_OUTPAD : MLP_HID;
    const int kl = l == 0 ? MLP_KPAD : MLP_HID;
    NtArgs a = nt_args(nout, kl, (int)rows);
    a.A = l == MLP_HIDDEN_LAYERS ? tt.dyT : tt.dT[l];
    a.lda = (int)ld;
    a.B = l == 0 ? tt.XT : tt.HT[l - 1];
    a.ldb = (int)ld;
    a.kps = kps;
    a.part = part;
    a.bpart = part + (size_t)slices * nout * kl;
    wg[l] = a;
    wsl[l] = slices;
    S.part[l] = a.part;
    S.bpart[l] = a.bpart;
    tt.part_w[l] = a.part;
    tt.part_b[l] = a.bpart;
    S.gw[l] = tb.gw[l];
    S.gb[l] = tb.gb[l];
    S.nw[l] = (uint32_t)(nout * kl);
    S.nb[l] = (uint32_t)nout;
    part += (size_t)slices * (nout * kl + nout);
  }
  if ((e = gemm_nt_multi(wg, wsl, st)) != hipSuccess) return e;
  tt.slices = slices;
  if (!reduce) return hipSuccess;
  uint32_t tot = 0;
  for (int l = 0; l < MLP_LAYERS; ++l) {
    S.first[l] = tot;
    tot += S.nw[l] + S.nb[l];
  }
  S.first[MLP_LAYERS] = tot;
  hipLaunchKernelGGL(k_grad_sum, dim3((tot + 255) / 256), dim3(256), 0, st, S);
  return hipGetLastError();
}

}  // namespace pm
