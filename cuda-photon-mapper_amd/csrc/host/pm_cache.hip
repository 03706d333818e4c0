// pm_cache.hip -- host side of the learned radiance cache: parameter
// initialisation, online training (Adam), inference and inspection, behind
// the pm_cache_* entry points of include/photon_mapper.h.
//
// No reference counterpart (the reference has no learned component).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "pm_context.h"

using namespace pm;

namespace {

const uint32_t kChunk = 1u << 20;  // rows per inference / training chunk

size_t w_offset(int l) {
  size_t o = 0;
  for (int i = 0; i < l; ++i) o += mlp_w_count(i) + mlp_b_count(i);
  return o;
}
size_t wb_offset(int l) {
  size_t o = 0;
  for (int i = 0; i < l; ++i) o += mlp_w_count(i);
  return o;
}

// Deterministic initialisation: He-uniform weights from Philox (seed, index),
// zero biases, zero padding (input columns >= 216, output rows >= 3).
std::vector<float> init_params(uint32_t seed) {
  std::vector<float> p(mlp_param_count(), 0.0f);
  for (int l = 0; l < MLP_LAYERS; ++l) {
    const int kin = mlp_kin(l), nout = mlp_nout(l);
    const int kreal = l == 0 ? MLP_FEAT : MLP_HID;
    const int nreal = l == MLP_HIDDEN_LAYERS ? MLP_OUT : MLP_HID;
    const float s = l == MLP_HIDDEN_LAYERS ? sqrtf(6.0f / (float)(kreal + nreal)) : sqrtf(6.0f / (float)kreal);
    float* w = p.data() + w_offset(l);
    for (int o = 0; o < nout; ++o)
      for (int i = 0; i < kin; ++i) {
        if (o >= nreal || i >= kreal) continue;
        uint32_t idx = (uint32_t)(w_offset(l) + (size_t)o * kin + i);
        float u = u01(rng(seed, idx, 0u, 0u, TAG_CACHE).x);
        w[(size_t)o * kin + i] = (2.0f * u - 1.0f) * s;
      }
  }
  return p;
}

int ensure_cache(pm_context* c) {
  CacheState& k = c->cache;
  if (k.ready) return PM_OK;
  const size_t np = mlp_param_count();
  size_t nwb = 0;
  for (int l = 0; l < MLP_LAYERS; ++l) nwb += mlp_w_count(l);
  HIPCHK(k.params.ensure(4 * np));
  HIPCHK(k.grads.ensure(4 * np));
  HIPCHK(k.adam_m.ensure(4 * np));
  HIPCHK(k.adam_v.ensure(4 * np));
  HIPCHK(k.params_bf16.ensure(2 * 2 * nwb));  // W, then W^T
  HIPCHK(k.params_ema.ensure(4 * np));
  HIPCHK(k.ema_bf16.ensure(2 * nwb));
  for (int l = 0; l < MLP_LAYERS; ++l) {
    k.w.w[l] = k.params.as<float>() + w_offset(l);
    k.w.b[l] = k.w.w[l] + mlp_w_count(l);
    k.w.mw[l] = k.adam_m.as<float>() + w_offset(l);
    k.w.mb[l] = k.w.mw[l] + mlp_w_count(l);
    k.w.vw[l] = k.adam_v.as<float>() + w_offset(l);
    k.w.vb[l] = k.w.vw[l] + mlp_w_count(l);
    k.w.wb[l] = k.params_bf16.as<uint16_t>() + wb_offset(l);
    k.w.wbt[l] = k.params_bf16.as<uint16_t>() + nwb + wb_offset(l);
    k.w.we[l] = k.params_ema.as<float>() + w_offset(l);
    k.w.be[l] = k.w.we[l] + mlp_w_count(l);
    k.w.wbe[l] = k.ema_bf16.as<uint16_t>() + wb_offset(l);
    k.tb.gw[l] = k.grads.as<float>() + w_offset(l);
    k.tb.gb[l] = k.tb.gw[l] + mlp_w_count(l);
  }
  const uint32_t R = kChunk;
  HIPCHK(k.X.ensure(2ull * R * MLP_KPAD));
  for (int l = 0; l < MLP_HIDDEN_LAYERS; ++l) HIPCHK(k.H[l].ensure(2ull * R * MLP_HID));
  HIPCHK(k.y.ensure(4ull * R * MLP_OUTPAD));
  HIPCHK(k.dy.ensure(2ull * R * MLP_OUTPAD));
  HIPCHK(k.row_err.ensure(4ull * R));
  k.tb.dy = k.dy.as<uint16_t>();
  k.tb.row_err = k.row_err.as<float>();
  k.rows_cap = R;
  k.ready = true;
  return PM_OK;
}

int upload_params(pm_context* c, const float* host) {
  CacheState& k = c->cache;
  HIPCHK(hipMemcpy(k.params.p, host, 4 * mlp_param_count(), hipMemcpyHostToDevice));
  HIPCHK(mlp_refresh_bf16(k.w, c->stream));
  HIPCHK(mlp_reset_ema(k.w, c->stream));
  k.ema_stale = false;
  HIPCHK(hipStreamSynchronize(c->stream));
  return PM_OK;
}

// The training pass's transposed activations and split partials for `rows`.
int ensure_train_t(CacheState& k, uint32_t rows, MlpTrainT& tt) {
  const uint32_t ld = mlp_train_ld(rows);
  HIPCHK(k.XT.ensure(2ull * MLP_KPAD * ld));
  for (int l = 0; l < MLP_HIDDEN_LAYERS; ++l) HIPCHK(k.HT[l].ensure(2ull * MLP_HID * ld));
  for (int j = 0; j < MLP_HIDDEN_LAYERS; ++j) HIPCHK(k.dT[j].ensure(2ull * MLP_HID * ld));
  HIPCHK(k.dyT.ensure(2ull * MLP_OUTPAD * ld));
  HIPCHK(k.tpart.ensure(4ull * mlp_train_partial_floats(rows)));
  tt.XT = k.XT.as<uint16_t>();
  for (int l = 0; l < MLP_HIDDEN_LAYERS; ++l) tt.HT[l] = k.HT[l].as<uint16_t>();
  for (int j = 0; j < MLP_HIDDEN_LAYERS; ++j) tt.dT[j] = k.dT[j].as<uint16_t>();
  tt.dyT = k.dyT.as<uint16_t>();
  tt.ld = ld;
  tt.part = k.tpart.as<float>();
  return PM_OK;
}

// The weights' moving average as the next use needs it: restarted at the
// current weights after the option changed.
int sync_ema(pm_context* c) {
  CacheState& k = c->cache;
  k.w.ema_decay = k.ema_decay;
  if (k.ema_stale && k.ema_decay > 0.0f) HIPCHK(mlp_reset_ema(k.w, c->stream));
  k.ema_stale = false;
  return PM_OK;
}

uint16_t* const* hidden_ptrs(CacheState& k, uint16_t** tmp) {
  for (int l = 0; l < MLP_HIDDEN_LAYERS; ++l) tmp[l] = k.H[l].as<uint16_t>();
  return tmp;
}

}  // namespace

namespace pm {
void cache_release(CacheState& k) {
  for (DevBuf* b : {&k.params, &k.grads, &k.adam_m, &k.adam_v, &k.params_bf16, &k.params_ema, &k.ema_bf16, &k.X, &k.y, &k.dy,
                    &k.row_err, &k.rgb, &k.err_sum, &k.err_cnt, &k.rec_cnt, &k.rec_off, &k.rec_in,
                    &k.rec_wt, &k.rec_code, &k.rec_val, &k.sel_in, &k.sel_t, &k.list, &k.list_cnt, &k.cell_fb,
                    &k.row_code, &k.row_list, &k.loss_part, &k.loss_acc, &k.rec_q, &k.map_hash, &k.list_fb,
                    &k.XT, &k.dT[0], &k.dT[1], &k.dT[2], &k.dT[3], &k.dyT, &k.tpart})
    b->release();
  for (int l = 0; l < MLP_HIDDEN_LAYERS; ++l) {
    k.H[l].release();
    k.HT[l].release();
  }
  k.ready = false;
}

// The two record lists of `rule` over N record codes, each in
// record order: list 0 at list[0, n[0]), list 1 at list[n[0], n[0] + n[1]).
// One count pass, a scan of the per-block counts, one scatter pass.
int build_lists(pm_context* c, const uint16_t* code, uint64_t N, const CacheListRule& rule, uint32_t* list,
                uint32_t n[2]) {
  CacheState& k = c->cache;
  hipStream_t st = c->stream;
  n[0] = n[1] = 0u;
  if (N == 0) return PM_OK;
  const size_t nb = cache_list_blocks(N);
  HIPCHK(k.list_cnt.ensure(4ull * 4 * nb));
  HIPCHK(c->scratch.ensure(4ull * scan_temp_words(2 * nb)));
  HIPCHK(c->small.ensure(64));
  uint32_t* counts = k.list_cnt.as<uint32_t>();
  uint32_t* offs = counts + 2 * nb;
  uint32_t* total = c->small.as<uint32_t>() + 12;
  HIPCHK(launch_cache_list_count(code, N, rule, counts, st));
  HIPCHK(exclusive_scan(counts, offs, 2 * nb, c->scratch.as<uint32_t>(), total, st));
  HIPCHK(launch_cache_list_scatter(code, N, rule, offs, list, st));
  uint32_t h[2] = {0u, 0u};  // list 0's total (the offset of list 1's first block), both totals
  HIPCHK(hipMemcpyAsync(&h[0], offs + nb, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&h[1], total, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  n[0] = h[0];
  n[1] = h[1] - h[0];
  return PM_OK;
}

// Encode + forward + decode for n device input rows, in chunks.  With `list`
// (nlist record indices, all of volume records) the fused kernel infers only
// the listed records; the other forms infer all n.
int cache_predict_dev(pm_context* c, const float* in, uint32_t n, float* rgb, const uint32_t* list = nullptr,
                      uint32_t nlist = 0) {
  int r = ensure_cache(c);
  if (r != PM_OK) return r;
  CacheState& k = c->cache;
  if ((r = sync_ema(c)) != PM_OK) return r;
  const MlpDev wi = mlp_inference_net(k.w);
  k.infer_rows = n;  // the fused list path below infers the listed rows only
  k.infer_short = 0;
  if (k.fused_infer) {
    int maxlds = 0;
    HIPCHK(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
    if (k.fused_infer >= 3 && (size_t)maxlds >= mlp_fused_lds_bytes()) {
      const bool split = k.fused_infer == 4;
      if (list) {
        HIPCHK(mlp_infer_fused(wi, in, nlist, rgb, c->stream, list, split));
        k.infer_rows = nlist;
        k.infer_short = split ? nlist : 0;
        return PM_OK;
      }
      if (!split || n == 0) {
        HIPCHK(mlp_infer_fused(wi, in, n, rgb, c->stream));
        return PM_OK;
      }
      // surface rows on the full layer 1, volume rows on the short one
      HIPCHK(k.row_code.ensure(2ull * n));
      HIPCHK(k.row_list.ensure(4ull * n));
      HIPCHK(launch_cache_row_codes(in, n, k.row_code.as<uint16_t>(), c->stream));
      CacheListRule kinds;
      memset(&kinds, 0, sizeof(kinds));
      uint32_t nk[2] = {0, 0};
      if ((r = build_lists(c, k.row_code.as<uint16_t>(), n, kinds, k.row_list.as<uint32_t>(), nk)) != PM_OK)
        return r;
      if (nk[0]) HIPCHK(mlp_infer_fused(wi, in, nk[0], rgb, c->stream, k.row_list.as<uint32_t>(), false));
      if (nk[1]) HIPCHK(mlp_infer_fused(wi, in, nk[1], rgb, c->stream, k.row_list.as<uint32_t>() + nk[0], true));
      k.infer_short = nk[1];
      return PM_OK;
    }
    if ((size_t)maxlds >= mlp_infer_lds_bytes()) {
      HIPCHK(mlp_infer(wi, in, n, k.H[0].as<uint16_t>(), rgb, c->stream, k.fused_infer == 2 ? 1 : 0,
                       1u << k.infer_chunk_log2));
      return PM_OK;
    }
  }
  uint16_t* Hp[MLP_HIDDEN_LAYERS];
  Hp[0] = k.H[0].as<uint16_t>();
  Hp[1] = k.H[1].as<uint16_t>();
  Hp[2] = k.H[0].as<uint16_t>();
  Hp[3] = k.H[1].as<uint16_t>();
  for (uint32_t o = 0; o < n; o += kChunk) {
    uint32_t rows = std::min(kChunk, n - o);
    HIPCHK(mlp_encode(in + (size_t)o * MLP_IN, rows, k.X.as<uint16_t>(), c->stream));
    HIPCHK(mlp_forward(wi, k.X.as<uint16_t>(), rows, Hp, k.y.as<float>(), c->stream));
    HIPCHK(mlp_decode(k.y.as<float>(), rows, rgb + (size_t)o * MLP_OUT, c->stream));
  }
  return PM_OK;
}

// One Adam step per `batch` rows of (inputs, targets); returns the mean loss.
// With loss_terms, the loss is not read back: its sum stays in k.loss_acc on
// the device and *loss_terms gets the number of terms (no host wait).
int cache_train_dev(pm_context* c, const float* in, const float* target, uint32_t n, uint32_t batch, float lr,
                    double* loss, unsigned long long* esum = nullptr, unsigned int* ecnt = nullptr,
                    uint64_t* loss_terms = nullptr) {
  int r = ensure_cache(c);
  if (r != PM_OK) return r;
  CacheState& k = c->cache;
  if (batch == 0 || batch > kChunk) return PM_ERR_INVALID;
  uint16_t* tmp[MLP_HIDDEN_LAYERS];
  uint16_t* const* H = hidden_ptrs(k, tmp);
  MlpTrainT tt;
  if ((r = ensure_train_t(k, batch, tt)) != PM_OK) return r;
  if ((r = sync_ema(c)) != PM_OK) return r;
  uint64_t cnt = 0;
  uint32_t ro = 0;  // rows of squared errors in k.row_err not yet summed
  HIPCHK(k.loss_acc.ensure(8));
  double* dloss = k.loss_acc.as<double>();
  HIPCHK(k.loss_part.ensure(8 * sum_rows_parts()));
  HIPCHK(hipMemsetAsync(dloss, 0, sizeof(double), c->stream));
  for (uint32_t o = 0; o + batch <= n; o += batch) {
    // each step's squared errors go to the next rows of k.row_err, summed in one pass when it is full
    if (ro + batch > k.rows_cap) {
      HIPCHK(launch_sum_rows(k.row_err.as<float>(), ro, dloss, k.loss_part.as<double>(), c->stream));
      ro = 0;
    }
    MlpTrainBufs tb = k.tb;
    tb.row_err = k.row_err.as<float>() + ro;
    HIPCHK(mlp_train_grads(k.w, in + (size_t)o * MLP_IN, batch, H, k.y.as<float>(),
                           target + (size_t)o * MLP_OUT, tb, tt, false, c->stream, esum, ecnt));
    k.step += 1;
    HIPCHK(mlp_adam(k.w, tb, lr, k.step, c->stream, &tt));
    ro += batch;
    cnt += (uint64_t)batch * MLP_OUT;
  }
  if (ro) HIPCHK(launch_sum_rows(k.row_err.as<float>(), ro, dloss, k.loss_part.as<double>(), c->stream));
  if (loss_terms) {
    *loss_terms = cnt;
    return PM_OK;
  }
  double total = 0.0;
  HIPCHK(hipMemcpyAsync(&total, dloss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (loss) *loss = cnt ? total / (double)cnt : 0.0;
  return PM_OK;
}
}  // namespace pm

extern "C" {

uint32_t pm_cache_param_count(void) { return (uint32_t)mlp_param_count(); }

int pm_cache_reset(pm_context* c, uint32_t seed) {
  if (!c) return PM_ERR_INVALID;
  int r = ensure_cache(c);
  if (r != PM_OK) return r;
  std::vector<float> p = init_params(seed);
  HIPCHK(hipMemset(c->cache.adam_m.p, 0, 4 * mlp_param_count()));
  HIPCHK(hipMemset(c->cache.adam_v.p, 0, 4 * mlp_param_count()));
  c->cache.step = 0;
  c->cache.frame = 0;
  return upload_params(c, p.data());
}

int pm_cache_params(pm_context* c, float* host, uint32_t count, int set) {
  if (!c || !host || count != mlp_param_count()) return PM_ERR_INVALID;
  int r = ensure_cache(c);
  if (r != PM_OK) return r;
  if (set) return upload_params(c, host);
  HIPCHK(hipMemcpy(host, c->cache.params.p, 4ull * count, hipMemcpyDeviceToHost));
  return PM_OK;
}

int pm_cache_grad(pm_context* c, const float* in_dev, const float* target_dev, uint32_t n, float* grad_host,
                  uint32_t count, double* loss) {
  if (!c || !in_dev || !target_dev || !grad_host || count != mlp_param_count()) return PM_ERR_INVALID;
  if (n == 0 || n > kChunk) return PM_ERR_INVALID;
  int r = ensure_cache(c);
  if (r != PM_OK) return r;
  CacheState& k = c->cache;
  uint16_t* tmp[MLP_HIDDEN_LAYERS];
  uint16_t* const* H = hidden_ptrs(k, tmp);
  MlpTrainT tt;
  if ((r = ensure_train_t(k, n, tt)) != PM_OK) return r;
  HIPCHK(mlp_train_grads(k.w, in_dev, n, H, k.y.as<float>(), target_dev, k.tb, tt, true,
                         c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(grad_host, k.grads.p, 4ull * count, hipMemcpyDeviceToHost));
  if (loss) {
    std::vector<float> err(n);
    HIPCHK(hipMemcpy(err.data(), k.row_err.p, 4ull * n, hipMemcpyDeviceToHost));
    double s = 0.0;
    for (uint32_t i = 0; i < n; ++i) s += err[i];
    *loss = s / ((double)n * MLP_OUT);
  }
  return PM_OK;
}

int pm_cache_train(pm_context* c, const float* in_dev, const float* target_dev, uint32_t n, uint32_t batch,
                   float lr, double* loss) {
  if (!c || !in_dev || !target_dev || !(lr > 0.0f)) return PM_ERR_INVALID;
  if (!c->cache.ready) {
    int r = pm_cache_reset(c, c->cfg.seed);
    if (r != PM_OK) return r;
  }
  int r = cache_train_dev(c, in_dev, target_dev, n, batch, lr, loss);
  if (r == PM_OK) HIPCHK(hipStreamSynchronize(c->stream));
  return r;
}

int pm_cache_predict(pm_context* c, const float* in_dev, uint32_t n, float* rgb_dev) {
  if (!c || (n && (!in_dev || !rgb_dev))) return PM_ERR_INVALID;
  if (!c->cache.ready) {
    int r = pm_cache_reset(c, c->cfg.seed);
    if (r != PM_OK) return r;
  }
  int r = cache_predict_dev(c, in_dev, n, rgb_dev);
  if (r == PM_OK) HIPCHK(hipStreamSynchronize(c->stream));
  return r;
}

int pm_cache_configure(pm_context* c, float threshold, uint32_t train_samples, uint32_t batch, float lr,
                       int march_stride, int cache_surface) {
  if (!c || !(threshold >= 0.0f) || batch == 0 || batch > kChunk || !(lr > 0.0f)) return PM_ERR_INVALID;
  if (train_samples % batch || march_stride < 1 || march_stride > 64) return PM_ERR_INVALID;
  c->cache.stride = march_stride;
  c->cache.cache_surface = cache_surface ? 1 : 0;
  c->cache.threshold = threshold;
  c->cache.train_samples = train_samples;
  c->cache.batch = batch;
  c->cache.lr = lr;
  return PM_OK;
}

int pm_cache_stats_get(pm_context* c, pm_cache_stats* out) {
  if (!c || !out) return PM_ERR_INVALID;
  *out = c->cache.st;
  return PM_OK;
}

static double ev_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.0f;
  return hipEventElapsedTime(&ms, a, b) == hipSuccess ? (double)ms : -1.0;
}

int pm_render_cache_crop(pm_context* c, int x0, int y0, int w, int h, float* rgb_dev, pm_uchar4* out_dev) {
  if (!c) return PM_ERR_INVALID;
  if (!c->built) return PM_ERR_STATE;
  if (w <= 0 || h <= 0 || x0 < 0 || y0 < 0 || x0 + w > c->cfg.width || y0 + h > c->cfg.height) return PM_ERR_INVALID;
  CacheState& k = c->cache;
  if (!k.ready) {
    int r = pm_cache_reset(c, c->cfg.seed);
    if (r != PM_OK) return r;
  }
  hipStream_t st = c->stream;
  hipEvent_t ev[7];
  for (int i = 0; i < 7; ++i) HIPCHK(hipEventCreate(&ev[i]));
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() {
      for (int i = 0; i < 7; ++i) (void)hipEventDestroy(e[i]);
    }
  } guard{ev};
  // record-order slots of the crop (4x4 pixel blocks, padded)
  const uint32_t npix = (uint32_t)cache_order_size(w, h);
  HIPCHK(hipEventRecord(ev[0], st));
  // 1-3: record the frame's queries in record order (4x4 pixel blocks)
  HIPCHK(k.rec_cnt.ensure(4ull * npix));
  HIPCHK(k.rec_off.ensure(4ull * npix));
  HIPCHK(c->scratch.ensure(4ull * scan_temp_words(npix)));
  CacheRecordLaunch rl;
  memset(&rl, 0, sizeof(rl));
  rl.sc = c->scene;
  rl.rp = c->rp;
  if (!c->render_media) rl.rp.media_on = 0;  // PM_OPT_RENDER_MEDIA 0: surface queries only
  rl.x0 = x0; rl.y0 = y0; rl.w = w; rl.h = h;
  rl.counts = k.rec_cnt.as<uint32_t>();
  rl.write = 0;
  rl.stride = k.stride;
  // near-light radius: near_groups merged groups (none without merging)
  const float near = k.stride > 1 ? (float)k.near_groups * (float)k.stride * c->rp.march_step : 0.0f;
  rl.near2 = near * near;
  HIPCHK(launch_cache_record(rl, st));
  HIPCHK(exclusive_scan(k.rec_cnt.as<uint32_t>(), k.rec_off.as<uint32_t>(), npix, c->scratch.as<uint32_t>(), nullptr,
                        st));
  // the record total in 64 bits: record offsets are 32-bit
  HIPCHK(c->small.ensure(64));
  unsigned long long* ntot_dev = (unsigned long long*)(c->small.as<uint32_t>() + 14);
  HIPCHK(launch_count_total(k.rec_cnt.as<uint32_t>(), npix, ntot_dev, st));
  uint64_t N = 0;
  HIPCHK(hipMemcpyAsync(&N, ntot_dev, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (N >= 0xffffffffull) return PM_ERR_CAPACITY;
  HIPCHK(k.rec_in.ensure(4ull * MLP_IN * std::max<uint64_t>(N, 1)));
  HIPCHK(k.rec_wt.ensure(12ull * std::max<uint64_t>(N, 1)));
  HIPCHK(k.rec_code.ensure(2ull * std::max<uint64_t>(N, 1)));
  HIPCHK(k.rec_val.ensure(12ull * std::max<uint64_t>(N, 1)));
  rl.write = 1;
  rl.offsets = k.rec_off.as<uint32_t>();
  rl.in = k.rec_in.as<float>();
  rl.wt = k.rec_wt.as<float>();
  rl.code = k.rec_code.as<uint16_t>();
  HIPCHK(launch_cache_record(rl, st));
  HIPCHK(hipEventRecord(ev[1], st));
  // 4: online training on a subset with exact kNN targets
  GatherMap gs = make_gather_map(c, c->ms.m, c->cfg.k_surface, c->cfg.rcap_surface);
  GatherMap gv = make_gather_map(c, c->mv.m, c->cfg.k_volume, c->cfg.rcap_volume);
  // exact values with the gather of mode 7 under modes 7 and 5 (auto)
  const int mode7 = c->gather_mode == 7 || c->gather_mode == 5 ? 1 : 0;
  const uint32_t ncell = 2u * CACHE_GRID * CACHE_GRID * CACHE_GRID;
  HIPCHK(k.err_sum.ensure(8ull * ncell));
  HIPCHK(k.err_cnt.ensure(4ull * ncell));
  HIPCHK(hipMemsetAsync(k.err_sum.p, 0, 8ull * ncell, st));
  HIPCHK(hipMemsetAsync(k.err_cnt.p, 0, 4ull * ncell, st));
  double loss = 0.0;
  uint64_t loss_terms = 0;  // the training loss's sum is read after the frame
  const uint32_t S = N > 0 ? k.train_samples : 0u;
  // the volume records, when surface records are never cached: inference
  // skips the surface records, whose values the exact fallback writes
  const uint32_t* infer_list = nullptr;
  uint32_t n_infer = 0;
  // seeds of the counting first pass: each fallback record's bucket from the
  // previous frame, kept while the records and the photon positions are
  // unchanged (the same record count and map fingerprints; a static scene
  // re-emitted with the same seed, or a reused map).  Any seed gives the same
  // values, but seeds from other photon positions miss often and cost more
  // than the heap passes they replace, so they are dropped.
  uint16_t* seeds = nullptr;
  bool seeds_checked = false;
  auto prepare_seeds = [&]() -> int {
    if (seeds_checked || !mode7 || !k.use_seeds) return PM_OK;
    seeds_checked = true;
    HIPCHK(k.map_hash.ensure(16));
    unsigned long long* mh = (unsigned long long*)k.map_hash.p;
    HIPCHK(launch_point_hash(c->ms.m.tpts, c->ms.m.n, mh, st));
    HIPCHK(launch_point_hash(c->mv.m.tpts, c->mv.m.n, mh + 1, st));
    unsigned long long hh[2] = {0ull, 0ull};
    HIPCHK(hipMemcpyAsync(hh, mh, 16, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (k.rec_q_n != N || !k.rec_q.p || hh[0] != k.rec_q_hash[0] || hh[1] != k.rec_q_hash[1]) {
      HIPCHK(k.rec_q.ensure(2ull * std::max<uint64_t>(N, 1)));
      HIPCHK(hipMemsetAsync(k.rec_q.p, 0xff, 2ull * N, st));
      k.rec_q_n = N;
      k.rec_q_hash[0] = hh[0];
      k.rec_q_hash[1] = hh[1];
    }
    seeds = k.rec_q.as<uint16_t>();
    return PM_OK;
  };
  // when every surface record falls back (surfaces not cached), their exact
  // values do not depend on the network: they are gathered on the second
  // stream while the network trains and infers, from the surface list of the
  // kind lists (k.list, not rewritten in this frame after that point)
  bool surface_early = false;
  if (S > 0) {
    HIPCHK(k.sel_in.ensure(4ull * MLP_IN * S));
    HIPCHK(k.sel_t.ensure(12ull * S));
    // index lists of surface and volume records (in record order)
    HIPCHK(k.list.ensure(8ull * std::max<uint64_t>(N, 1)));
    CacheListRule kinds;
    memset(&kinds, 0, sizeof(kinds));
    uint32_t nk[2] = {0, 0};
    int rl_err = build_lists(c, k.rec_code.as<uint16_t>(), N, kinds, k.list.as<uint32_t>(), nk);
    if (rl_err != PM_OK) return rl_err;
    const uint32_t* list_s = k.list.as<uint32_t>();
    const uint32_t* list_v = list_s + nk[0];
    if (!k.cache_surface) {
      infer_list = list_v;
      n_infer = nk[1];
    }
    if (mode7 && !k.cache_surface && nk[0] > 0) {
      int rs = prepare_seeds();
      if (rs != PM_OK) return rs;
      HIPCHK(hipEventRecord(c->ev_fork, st));
      HIPCHK(hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
      CacheValueLaunch vs;
      memset(&vs, 0, sizeof(vs));
      vs.gs = gs; vs.gv = gv; vs.hg_g = c->scene.med.g;
      vs.in = k.rec_in.as<float>();
      vs.list = list_s;
      vs.n = nk[0];
      vs.out = k.rec_val.as<float>();
      vs.out_by_list = 1;
      vs.mode7 = mode7;
      vs.kind = 1;
      vs.seed_q = seeds;
      HIPCHK(launch_cache_values(vs, c->stream2));
      HIPCHK(hipEventRecord(c->ev_join, c->stream2));
      surface_early = true;
    }
    // without surface caching the subset is drawn from the volume records only
    const uint32_t ns_sel = k.cache_surface ? nk[0] : 0u;
    if (ns_sel > 0 || nk[1] > 0) {  // else there is nothing the network answers, and nothing to train on
      HIPCHK(launch_cache_select(k.rec_in.as<float>(), list_s, ns_sel, list_v, nk[1], S, c->cfg.seed, k.frame,
                                 k.sel_in.as<float>(), st));
      CacheValueLaunch vl;
      memset(&vl, 0, sizeof(vl));
      vl.gs = gs; vl.gv = gv; vl.hg_g = c->scene.med.g;
      vl.in = k.sel_in.as<float>();
      vl.n = S;
      vl.out = k.sel_t.as<float>();
      vl.mode7 = mode7;
      HIPCHK(launch_cache_values(vl, st));
      int r = cache_train_dev(c, k.sel_in.as<float>(), k.sel_t.as<float>(), S, k.batch, k.lr, nullptr,
                              k.err_sum.as<unsigned long long>(), k.err_cnt.as<unsigned int>(), &loss_terms);
      if (r != PM_OK) return r;
    }
  }
  HIPCHK(hipEventRecord(ev[2], st));
  // 5: inference over the records the cache may answer
  int r = cache_predict_dev(c, k.rec_in.as<float>(), (uint32_t)N, k.rec_val.as<float>(), infer_list, n_infer);
  if (r != PM_OK) return r;
  HIPCHK(hipEventRecord(ev[3], st));
  // 6: exact fallback where the cache's cell error is high.  The list holds
  // the surface records, then the volume records, each in record order: a
  // wave then runs one kind of gather
  HIPCHK(k.list.ensure(8ull * std::max<uint64_t>(N, 1)));
  HIPCHK(k.cell_fb.ensure(CACHE_CELLS));
  HIPCHK(launch_cache_cell_fb(k.err_sum.as<unsigned long long>(), k.err_cnt.as<unsigned int>(), k.threshold,
                              k.cell_fb.as<uint8_t>(), st));
  CacheListRule fbr;
  fbr.fallback = 1;
  fbr.cache_surface = k.cache_surface;
  fbr.cell_fb = k.cell_fb.as<uint8_t>();
  uint32_t nk2[2] = {0, 0};
  HIPCHK(k.list_fb.ensure(8ull * std::max<uint64_t>(N, 1)));
  r = build_lists(c, k.rec_code.as<uint16_t>(), N, fbr, k.list_fb.as<uint32_t>(), nk2);
  if (r != PM_OK) return r;
  const uint32_t nfb = nk2[0] + nk2[1];
  if (mode7 && nfb > 0) {
    r = prepare_seeds();
    if (r != PM_OK) return r;
  }
  for (int kind = 1; kind <= 2; ++kind) {  // one launch per kind of row
    if (nk2[kind - 1] == 0 || (kind == 1 && surface_early)) continue;
    CacheValueLaunch vl;
    memset(&vl, 0, sizeof(vl));
    vl.gs = gs; vl.gv = gv; vl.hg_g = c->scene.med.g;
    vl.in = k.rec_in.as<float>();
    vl.list = k.list_fb.as<uint32_t>() + (kind == 2 ? nk2[0] : 0u);
    vl.n = nk2[kind - 1];
    vl.out = k.rec_val.as<float>();
    vl.out_by_list = 1;
    vl.mode7 = mode7;
    vl.kind = kind;
    vl.seed_q = seeds;
    HIPCHK(launch_cache_values(vl, st));
  }
  if (surface_early) HIPCHK(hipStreamWaitEvent(st, c->ev_join, 0));
  HIPCHK(hipEventRecord(ev[4], st));
  // 7: per-pixel accumulation
  HIPCHK(launch_cache_accumulate(k.rec_off.as<uint32_t>(), k.rec_cnt.as<uint32_t>(), w, h, k.rec_wt.as<float>(),
                                 k.rec_val.as<float>(), c->rp.spp, c->cfg.exposure, rgb_dev, (uchar4*)out_dev, st));
  HIPCHK(hipEventRecord(ev[5], st));
  HIPCHK(hipStreamSynchronize(st));
  if (loss_terms) {
    double total = 0.0;
    HIPCHK(hipMemcpy(&total, k.loss_acc.p, sizeof(double), hipMemcpyDeviceToHost));
    loss = total / (double)loss_terms;
  }
  k.frame += 1;
  k.st.records = N;
  k.st.fallbacks = nfb;
  k.st.train_loss = loss;
  k.st.ms_record = ev_ms(ev[0], ev[1]);
  k.st.ms_train = ev_ms(ev[1], ev[2]);
  k.st.ms_infer = ev_ms(ev[2], ev[3]);
  k.st.ms_fallback = ev_ms(ev[3], ev[4]);
  k.st.ms_accumulate = ev_ms(ev[4], ev[5]);
  k.st.ms_total = ev_ms(ev[0], ev[5]);
  k.st.frame = k.frame;
  k.st.inferred = k.infer_rows;
  k.st.inferred_volume = (uint32_t)k.infer_short;
  return PM_OK;
}

}  // extern "C"
