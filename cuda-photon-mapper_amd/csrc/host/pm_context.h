// pm_context.h -- the library context behind the C ABI (internal to libpm.so).
#pragma once

#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../../include/photon_mapper.h"
#include "../core/pm_presets.h"
#include "../kernels/pm_kernels.h"
#include "../kernels/pm_mlp.h"

struct pm_context;

namespace pm {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      p = nullptr;
      bytes = 0;
      if (e != hipSuccess) return e;
    }
    size_t want = b ? b : 16;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) bytes = want;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T* as() const { return (T*)p; }
};

struct MapBufs {
  DevBuf x, y, z, payload, tpts, nodes, nrange, bmin, bmax, tpayload;
  DevBuf shape;                         // the kd-tree's shape tables for (shape_n, shape_L)
  uint32_t shape_n = 0, shape_L = 0;
  DevMap m;
  void release() {
    x.release(); y.release(); z.release(); payload.release(); tpts.release(); nodes.release(); nrange.release();
    bmin.release(); bmax.release(); tpayload.release(); shape.release();
    shape_n = shape_L = 0;
  }
};

// Learned radiance cache state: parameters, Adam moments, training scratch.
struct CacheState {
  DevBuf params, grads, adam_m, adam_v, params_bf16, params_ema, ema_bf16;
  DevBuf X, H[MLP_HIDDEN_LAYERS], y, dy, row_err, rgb;
  DevBuf XT, HT[MLP_HIDDEN_LAYERS], dT[MLP_HIDDEN_LAYERS], dyT, tpart;  // training: transposes and split partials
  MlpDev w;
  MlpTrainBufs tb;
  uint32_t rows_cap = 0;
  int step = 0;
  bool ready = false;
  // per-cell error tracker (fallback to the plain gather where it is high)
  DevBuf err_sum, err_cnt;
  // cache render configuration and per-frame buffers
  float threshold = 0.05f;  // per-channel mean squared log error of a cell
  uint32_t train_samples = 1u << 18;
  uint32_t batch = 1u << 14;
  float lr = 1e-3f;
  uint32_t frame = 0;
  int fused_infer = 4;     // pm_mlp_infer.hip kernels (when the LDS fits): 4 = one fused kernel, activations in
                           // registers, volume records on the 5-step layer 1; 3 = the same kernel, full layer 1
                           // for every row; 2 = layer 1 + layers 2..5 on 16x16x32 MFMA, 1 = 16x16x16; 0: generic GEMMs
  DevBuf row_code, row_list;  // pm_cache_predict under mode 4: the rows' codes and kind lists
  DevBuf loss_part, loss_acc;  // partial sums of the training loss, and its sum
  DevBuf rec_q;               // per record, the k-th neighbour's bucket of its last exact gather (seeds)
  uint64_t rec_q_n = 0;       // the record count rec_q was kept for
  int use_seeds = 1;          // PM_OPT_CACHE_SEEDS
  DevBuf map_hash;            // fingerprints of the two maps' point positions (this frame)
  unsigned long long rec_q_hash[2] = {0ull, 0ull};  // ... and of the maps rec_q was kept for
  uint64_t infer_rows = 0;    // rows of the last inference, and of them on the 5-step layer 1
  uint64_t infer_short = 0;
  int infer_chunk_log2 = 20;  // rows per fused-inference launch pair (log2)
  int stride = 1;          // march points per volume record
  int cache_surface = 0;   // 0: surface queries are always gathered exactly
  int near_groups = 2;     // near-light radius in merged march groups
  float ema_decay = 0.95f;  // inference on the moving average of the weights (decay per Adam step), 0: off
  bool ema_stale = true;    // the average must restart at the current weights before its next use
  DevBuf rec_cnt, rec_off, rec_in, rec_wt, rec_code, rec_val, sel_in, sel_t, list, list_cnt, cell_fb;
  DevBuf list_fb;  // the fallback lists (list keeps the kind lists, which the early surface gather reads)
  pm_cache_stats st;
};

void set_last_hip(hipError_t e, const char* what);
GatherMap make_gather_map(const pm_context* c, const DevMap& m, int k, float rcap);
void shape_tables(uint32_t n, uint32_t L, std::vector<uint32_t>& lo, std::vector<uint32_t>& hi,
                  std::vector<uint32_t>& lc, std::vector<uint32_t>& lb);
// Host copies of a map's records / tree (the layouts of pm_get_records / pm_get_tree).
int records_to_host(const DevMap& m, float* pos, float* power, float* dir);
int tree_to_host(const DevMap& m, uint32_t* perm, float* split, uint32_t* axis);
extern char g_last_err[256];

#define HIPCHK(expr)                          \
  do {                                        \
    hipError_t _e = (expr);                   \
    if (_e != hipSuccess) {                   \
      pm::set_last_hip(_e, #expr);            \
      return PM_ERR_HIP;                      \
    }                                         \
  } while (0)

}  // namespace pm

struct pm_context {
  pm_config cfg;
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // the volume map's kd-build, beside the surface map's on `stream`
  std::vector<pm::Sphere> spheres_h;
  pm::DevBuf spheres_d, density_d;
  pm::Scene scene;
  pm::TraceParams tp;
  pm::RenderParams rp;
  pm::MapBufs ms, mv;
  pm::DevBuf cnt_s, cnt_v, off_s, off_v, scratch, scratch2, tally_part, tally9, small, counters;
  uint32_t n_emit_total = 0;
  uint32_t caps_N = 0;  // caps mode: the photons the last emission needed to meet both caps (its N_total)
  double tally_h[9];
  bool emitted = false, built = false;
  int gather_mode = 5;  // auto
  int blocks_per_cu = 4;
  int tile_regions = 0;  // per-XCD tile regions of the lane render (0: raster order)
  int sample_lanes = 1;  // mode 7: a pixel's samples in adjacent lanes
  int render_media = 1;  // 0: the plain render gathers at surface hits only
  int top_levels = 0;    // mode 7: kd-tree levels staged in LDS per wave
  int voxel_grid = 0;    // 1: the plain render uses the voxel-grid estimate, 2: with the bilinear blend
  bool vox_ready = false;  // voxel grids and tables hold the current records
  pm::DevBuf vox_grid_s, vox_grid_v, vox_sat_s, vox_sat_v;
  pm_stats stats;
  hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;  // the two-stream kd-build's ordering
  pm::CacheState cache;
};
