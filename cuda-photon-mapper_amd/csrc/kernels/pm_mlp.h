// pm_mlp.h -- dimensions and device buffers of the learned radiance cache.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>

namespace pm {

enum : int {
  MLP_IN = 9,                         // position, direction, normal (or 0)
  MLP_BANDS = 12,                     // sin/cos frequency bands per input
  MLP_FEAT = MLP_IN * 2 * MLP_BANDS,  // 216
  MLP_KPAD = 224,                     // features padded to a multiple of 32
  MLP_HID = 128,
  MLP_HIDDEN_LAYERS = 4,
  MLP_OUT = 3,
  MLP_OUTPAD = 16,                    // output rows padded to one MFMA tile
  MLP_LAYERS = MLP_HIDDEN_LAYERS + 1,
};

// Weight matrix of layer l is [nout x kin], row-major.
inline int mlp_kin(int l) { return l == 0 ? MLP_KPAD : MLP_HID; }
inline int mlp_nout(int l) { return l == MLP_HIDDEN_LAYERS ? MLP_OUTPAD : MLP_HID; }
inline size_t mlp_w_count(int l) { return (size_t)mlp_nout(l) * (size_t)mlp_kin(l); }
inline size_t mlp_b_count(int l) { return (size_t)mlp_nout(l); }
inline size_t mlp_param_count() {
  size_t n = 0;
  for (int l = 0; l < MLP_LAYERS; ++l) n += mlp_w_count(l) + mlp_b_count(l);
  return n;
}
// Multiply-adds per forward sample (unpadded sizes): 216*128 + 3*128*128 + 128*3.
inline double mlp_forward_macs() { return 216.0 * 128 + 3.0 * 128 * 128 + 128.0 * 3; }

struct MlpDev {
  float* w[MLP_LAYERS];       // FP32 master weights
  float* b[MLP_LAYERS];
  uint16_t* wb[MLP_LAYERS];   // BF16 copies used by the MFMA GEMMs
  uint16_t* wbt[MLP_LAYERS];  // BF16 transposes [kin x nout] (the training's delta products)
  float* mw[MLP_LAYERS];      // Adam moments
  float* vw[MLP_LAYERS];
  float* mb[MLP_LAYERS];
  float* vb[MLP_LAYERS];
  // exponential moving average of the weights (the inference network when
  // ema_decay > 0): FP32 weights and biases, BF16 weights
  float* we[MLP_LAYERS];
  float* be[MLP_LAYERS];
  uint16_t* wbe[MLP_LAYERS];
  float ema_decay;            // per Adam step; 0: no average
};

struct MlpTrainBufs {
  uint16_t* dy;       // [rows x OUTPAD]
  float* row_err;     // [rows]
  float* gw[MLP_LAYERS];
  float* gb[MLP_LAYERS];
};

// Transposed copies kept by the training pass (pm_mlp_train.hip): every
// activation and delta also as [features x ld] (ld = rows rounded up to 64),
// the operands of the weight-gradient products, and the split partials.
struct MlpTrainT {
  uint16_t* XT;                     // [KPAD x ld]
  uint16_t* HT[MLP_HIDDEN_LAYERS];  // [HID x ld]
  uint16_t* dT[MLP_HIDDEN_LAYERS];  // [HID x ld]: dT[l] = the delta of hidden layer l's output, transposed
  uint16_t* dyT;                    // [OUTPAD x ld]
  uint32_t ld;
  float* part;                      // mlp_train_partial_floats(rows)
  // filled by mlp_train_grads: layer l's weight and bias partials, slice z at
  // part_w[l] + z * (its weight count) and part_b[l] + z * (its bias count)
  const float* part_w[MLP_LAYERS];
  const float* part_b[MLP_LAYERS];
  int slices;
};
int mlp_train_slices(uint32_t rows);
uint32_t mlp_train_ld(uint32_t rows);
size_t mlp_train_partial_floats(uint32_t rows);
// Encode, forward (keeping H and its transposes), loss, backward for `rows`
// inputs / targets: y the outputs, tb.row_err the per-row squared errors, and
// the gradients as split partials (tt.part_w / part_b); with `reduce` they are
// also summed into tb.gw / tb.gb (mlp_adam can instead sum them itself).
// With esum / ecnt, each row's squared error (capped at 1e6, in 2^-24 fixed
// point) is also added to the error tracker's cell of its input (err_cell),
// and the cell's row count incremented.
hipError_t mlp_train_grads(const MlpDev& w, const float* in, uint32_t rows, uint16_t* const* H,
                           float* y, const float* target, MlpTrainBufs& tb, MlpTrainT& tt, bool reduce,
                           hipStream_t st, unsigned long long* esum = nullptr, unsigned int* ecnt = nullptr);

hipError_t mlp_encode(const float* in, uint32_t n, uint16_t* out, hipStream_t st);
hipError_t mlp_forward(const MlpDev& w, const uint16_t* X, uint32_t rows, uint16_t* const* H, float* y,
                       hipStream_t st);
hipError_t mlp_decode(const float* y, uint32_t rows, float* rgb, hipStream_t st);
// One Adam step on the gradients in tb.gw / tb.gb, or, with parts, on the sum
// over slices (in slice order) of the partials mlp_train_grads left there.
hipError_t mlp_adam(MlpDev& w, const MlpTrainBufs& tb, float lr, int step, hipStream_t st,
                    const MlpTrainT* parts = nullptr);
hipError_t mlp_refresh_bf16(MlpDev& w, hipStream_t st);
// The moving average restarted at the current weights (FP32 and BF16 copies).
hipError_t mlp_reset_ema(MlpDev& w, hipStream_t st);
// The network inference reads: the moving average when ema_decay > 0, else w.
MlpDev mlp_inference_net(const MlpDev& w);
// Fused inference (pm_mlp_infer.hip): h1 holds mlp_infer_h1_rows() x 128 BF16.
size_t mlp_infer_h1_rows();
size_t mlp_infer_lds_bytes();
// k32 = 1: the v_mfma_f32_16x16x32_bf16 kernels, 0: the 16x16x16 ones;
// rows per launch pair (layer 1, layers 2..5) = min(chunk_rows, mlp_infer_h1_rows()).
hipError_t mlp_infer(const MlpDev& w, const float* in, uint32_t n, uint16_t* h1, float* rgb, hipStream_t st,
                     int k32, uint32_t chunk_rows);
// All five layers in one kernel with the activations in registers (pm_mlp_infer.hip).
// With `list`, row i is record list[i] of `in` and `rgb` (n = the list's
// length); the other records' outputs are left as they are.  `volume`: every
// row is a volume record (inputs 6..8 zero), which takes a shorter layer 1.
size_t mlp_fused_lds_bytes();
hipError_t mlp_infer_fused(const MlpDev& w, const float* in, uint32_t n, float* rgb, hipStream_t st,
                           const uint32_t* list = nullptr, bool volume = false);

}  // namespace pm
