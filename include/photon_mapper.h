/* photon_mapper.h -- C ABI of the MI355X photon mapper (libpm.so).
 *
 * Drop-in boundary for the reference's photonMappingKernel.cu launch ABI
 * (declared by its caller at simplePBO.cpp:27-29, defined at
 * photonMappingKernel.cu:1523-1582), plus an additive status-returning pm_*
 * API.  Every pointer named *_dev is a device pointer owned by the caller;
 * the library never frees or retains it.  Photon-map state persists inside a
 * library-owned context between calls (the reference keeps it in __device__
 * globals, photonMappingKernel.cu:76,80).
 *
 * Threading: calls on one context are not re-entrant; every call is
 * synchronous (returns after the GPU work finished), like the reference's
 * launchers (cudaThreadSynchronize after each launch).
 */
#ifndef PHOTON_MAPPER_H
#define PHOTON_MAPPER_H

#include <stdbool.h> /* the reference launchers' bool flags, also for C (cgo) callers */
#include <stddef.h>
#include <stdint.h>

/* uchar4 {x=r, y=g, z=b, w=0}: same 4-byte layout as HIP's/CUDA's uchar4. */
#ifndef PM_UCHAR4_DEFINED
#define PM_UCHAR4_DEFINED
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
typedef uchar4 pm_uchar4;
#else
typedef struct pm_uchar4 { unsigned char x, y, z, w; } pm_uchar4;
#endif
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ legacy ABI
 * Same symbols, signatures and error behaviour (print + exit(EXIT_FAILURE)) as
 * the reference.  They drive a default context on the current HIP device
 * configured as the unit-room scene (100k photons, k=32, r<=0.05, 1 spp).
 */

/* Replaces photonMappingKernel.cu:1569-1582 (fills the MWC random table).
 * Here: creates (or resets) the default context; Philox needs no table. */
void launch_init_random_numbers_kernel(void);

/* Replaces photonMappingKernel.cu:1523-1546 (grid reset + photon emission).
 * Here: traces photons and builds the kd-trees.  pos/width/height are unused
 * (as in the reference).  animTime moves the two spheres (reference
 * positionObjects, :1380-1404); participatingMediaFlag enables the medium
 * (:1239); interpolateFlag selects the cone-filtered estimate (the
 * reference's bilinear window blend, :591). */
void launch_emit_photons_kernel(pm_uchar4* pos, unsigned int image_width, unsigned int image_height,
                                float animTime, bool interpolateFlag, bool participatingMediaFlag);

/* Replaces photonMappingKernel.cu:1549-1566 (one thread per pixel into the
 * GL PBO).  Here: renders image_width x image_height into the caller's
 * device buffer `pos` (W*H uchar4), 1 sample per pixel, from the photon maps
 * of the last launch_emit_photons_kernel call (the first render emits when
 * none was made).  A new animTime or participatingMediaFlag changes the scene
 * the eye rays see at once; the photon maps change at the next emission, as
 * the reference's grid does while its emitFlag is off (callbacksPBO.cpp:60-61). */
void launch_photon_mapping_kernel(pm_uchar4* pos, unsigned int image_width, unsigned int image_height,
                                  float animTime, bool interpolateFlag, bool participatingMediaFlag);

/* The context behind the three calls above (created by the first of them
 * when needed), so that a legacy host can set options on it, e.g.
 * pm_set_option(pm_legacy_context(), PM_OPT_VOXEL_GRID, 1) for the
 * reference's own 32^3 voxel-grid estimate; with that option interpolateFlag
 * selects the bilinear window blend (photonMappingKernel.cu:591) instead of
 * the cone filter.  NULL when the context cannot be created.  No reference
 * counterpart: the reference chooses its estimate at compile time. */
struct pm_context* pm_legacy_context(void);

/* ------------------------------------------------------------------ pm_* API
 * Additive API (no reference counterpart; the reference configures itself
 * through compile-time macros at photonMappingKernel.cu:7-46 and __device__
 * globals at :59-98).  All functions return PM_OK or a negative status. */

enum {
  PM_OK = 0,
  PM_ERR_INVALID = -1,   /* null pointer, bad size or out-of-range parameter */
  PM_ERR_HIP = -2,       /* a HIP runtime call failed */
  PM_ERR_STATE = -3,     /* call order violated (e.g. render before emit) */
  PM_ERR_CAPACITY = -4,  /* a bounded buffer would overflow (nothing was written) */
  PM_ERR_NOMEM = -5
};

#define PM_MAX_SPHERES 256

enum { PM_MAT_DIFFUSE = 0, PM_MAT_MIRROR = 1, PM_MAT_DIELECTRIC = 2 };
enum { PM_MEDIUM_NONE = 0, PM_MEDIUM_HOMOGENEOUS = 1, PM_MEDIUM_NOISE = 2 };
enum { PM_STORE_FP32 = 0, PM_STORE_BF16 = 1 };
enum { PM_EST_BOX = 0, PM_EST_CONE = 1 };
enum { PM_MAP_SURFACE = 0, PM_MAP_VOLUME = 1 };

typedef struct pm_sphere {
  float center[3];
  float radius;
  float albedo[3];
  int material;
  float ior;
} pm_sphere;

typedef struct pm_config {
  uint32_t seed;
  /* scene: unit room [0,1]^3, walls 0:x=1 1:y=0 2:x=0 3:y=1 4:z=1, front open */
  float wall_albedo[5][3];
  int n_spheres;
  pm_sphere spheres[PM_MAX_SPHERES];
  float light_pos[3];
  float light_power[3]; /* W per channel */
  /* medium */
  int medium;
  float sigma_s, sigma_a, hg_g;
  int noise_res, noise_freq;
  /* photons */
  uint32_t n_emit;       /* photons emitted; 0 = emit until both caps are filled */
  uint32_t cap_surface;  /* record caps (0 = unbounded when n_emit > 0) */
  uint32_t cap_volume;
  int max_bounces;
  float rr_survive;
  int leaf_size;
  int storage;           /* PM_STORE_FP32 or PM_STORE_BF16 (power and direction) */
  /* gather */
  int k_surface, k_volume;
  float rcap_surface, rcap_volume;
  int estimator;
  /* render */
  int width, height, spp;
  float march_step;
  int max_depth;
  float eye[3];
  float focal;
  float exposure;
  int fp64_transmittance;
  int media_on;
  float anim_time;
} pm_config;

typedef struct pm_stats {
  uint32_t n_surface, n_volume;              /* records in each map */
  uint32_t n_emit_surface, n_emit_volume;    /* photons emitted for each map's normalisation */
  uint32_t n_emit_total;                     /* photons traced */
  uint32_t pad0;
  double emitted[3], absorbed[3], escaped[3], rr_gain[3]; /* FP64 flux tally, watts */
  double ms_emit, ms_build, ms_render;       /* HIP-event times of the last calls */
  double ms_gather_kernel;                   /* render kernel alone */
  uint64_t surface_queries, volume_queries;  /* radiance queries of the last render */
} pm_stats;

typedef struct pm_context pm_context;

/* Presets: 1 = unit room, 100k photons, homogeneous medium, k=32, 256x256;
 * 2 = 1M surface + 2M volume records, value-noise medium, k=64, 1920x1080x4;
 * 3 = preset 1 with BF16 power/direction and FP64 transmittance, 480x270;
 * 5 = preset 1 scene for the learned radiance cache. */
int pm_config_preset(int preset, pm_config* out);

int pm_create(const pm_config* cfg, pm_context** out);
int pm_destroy(pm_context* ctx);
/* Trace photons and fill both photon maps (records in (photon id, event) order). */
int pm_emit(pm_context* ctx);
/* Build the kd-trees of both maps. */
int pm_build(pm_context* ctx);
/* Render the configured frame (or the crop x0,y0,w,h when w>0): rgb_dev receives
 * w*h*3 FP32 radiance (may be NULL), out_dev w*h uchar4 (may be NULL). */
int pm_render(pm_context* ctx, float* rgb_dev, pm_uchar4* out_dev);
int pm_render_crop(pm_context* ctx, int x0, int y0, int w, int h, float* rgb_dev, pm_uchar4* out_dev);
/* pm_render_crop without any allocation: no reference counterpart.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops).  Same arguments, outputs and statistics as
 * pm_render_crop.  The render's only scratch, its two 64-bit query counters
 * (16 bytes), lives in a library-wide device array, one per GPU, held from
 * library load to process end; the call clears it with a kernel before the
 * render.  Calls that overlap on one GPU (several host threads or contexts)
 * share that array and race on it.  The context must already be built (its
 * photon maps and trees are allocated by pm_emit/pm_build). */
int pm_render_crop_noalloc(pm_context* ctx, int x0, int y0, int w, int h, float* rgb_dev, pm_uchar4* out_dev);
int pm_stats_get(pm_context* ctx, pm_stats* out);
/* Change the per-render parameters (resolution, spp, k, caps) without re-emitting. */
int pm_set_render(pm_context* ctx, int width, int height, int spp, int k_surface, int k_volume,
                  float rcap_surface, float rcap_volume);

/* Options.  PM_OPT_RENDER_MEDIA and PM_OPT_VOXEL_GRID change what is
 * estimated; the others change only how the GPU runs it, never the results:
 * PM_OPT_GATHER_MODE   0 = one wave64 per radiance query (cross-lane sorted list),
 *                      1 = one lane per query, 8x8 pixel tile per wave,
 *                          8-byte (d^2, index) heap keys in LDS,
 *                      2 = as 1 with 4-byte heap keys (d^2 only) and a second
 *                          fixed-radius pass for the exact neighbour set,
 *                      4 = as 2 with 2-byte keys (d^2 buckets 2^-7 wide),
 *                          the exact order restored inside the k-th bucket,
 *                          leaf points loaded 4 at a time under a
 *                          5-waves-per-SIMD register budget,
 *                      5 = auto (default): 7 when the frame has at least 16
 *                          8x8 tiles per compute unit and the maps hold at
 *                          least 2^20 records, else 1;
 *                      7 = as 4, with each march step's second pass merged
 *                          into the next step's first-pass traversal and the
 *                          first passes done by counting in a window of 256
 *                          buckets around the previous step's
 *                          k-th-neighbour bucket (heap pass on a miss);
 *                      (3 and 6 are not used: PM_ERR_INVALID);
 *                      every mode gives the same neighbour sets and images;
 * PM_OPT_BLOCKS_PER_CU workgroups per compute unit of the persistent mode-0 grid;
 * PM_OPT_CACHE_FUSED_INFER 4 = cache inference through one fused kernel, all
 *                      five layers with the activations in registers and the
 *                      weights in LDS, volume records (zero normal) through a
 *                      layer 1 without the normal's constant features (5 K
 *                      steps of 32 instead of 7; default), 3 = the same kernel
 *                      with the full layer 1 for every row, 2 = two LDS-resident kernels
 *                      (layer 1, layers 2..5) with 16x16x32 BF16 MFMA, 1 = the
 *                      same with 16x16x16 BF16 MFMA, 0 = the generic GEMMs;
 * PM_OPT_CACHE_INFER_CHUNK log2 of the rows per fused-inference launch pair,
 *                      14..20 (default 20): smaller chunks keep the layer-1
 *                      activations in the on-die caches between the two launches,
 *                      larger ones launch fewer, fuller grids (faster at 1080p);
 * PM_OPT_TILE_REGIONS  order in which the lane modes' workgroups take 8x8 tiles:
 *                      0 = raster order (default), rx = 1, 2, 4 or 8 = the
 *                      frame cut into rx x (8 / rx) regions, one per XCD
 *                      (workgroups go to the 8 XCDs round-robin), so that each
 *                      XCD's L2 serves one compact part of the frame;
 * PM_OPT_SAMPLE_LANES  gather mode 7 with spp = 2, 4, 8 or 16: 1 = the samples of a
 *                      pixel in adjacent lanes (default; a wave renders 64 / spp
 *                      pixels in raster tile order), 0 = one lane loops over a
 *                      pixel's samples (PM_OPT_TILE_REGIONS applies to this form);
 * PM_OPT_RENDER_MEDIA  1 = pm_render / pm_render_crop / pm_render_cache_crop
 *                      march the medium (default, when the scene has one),
 *                      0 = they gather at surface hits only: the surface-gather
 *                      measurement on the same photon maps (the volume map is
 *                      built but not read);
 * PM_OPT_CACHE_NEAR_GROUPS the cache render's near-light radius, in merged
 *                      march groups (0..16, default 2; a group is the march
 *                      stride's worth of steps): inside it every march point
 *                      is its own query and is gathered exactly;
 * PM_OPT_CACHE_EMA     the cache's inference network: 0 = the trained weights,
 *                      d = 1..999 = their exponential moving average with
 *                      decay d / 1000 per Adam step (default 950; training
 *                      itself is unchanged; setting the parameters or a reset
 *                      restarts the average);
 * PM_OPT_CACHE_SEEDS   1 = the cache render's exact fallback (gather modes 7
 *                      and auto) starts each record from its k-th neighbour's
 *                      bucket of the previous call while the records and the
 *                      photon positions are unchanged (default), 0 = heap
 *                      passes every call; the values are the same bits;
 * PM_OPT_TOP_LEVELS    gather mode 7 (also when auto picks it): the top 0..7
 *                      levels of both kd-trees (2^levels - 1 nodes, 48 B each)
 *                      staged in each wave's LDS for the traversals to read,
 *                      0 = none (default), every node read from HBM;
 * PM_OPT_VOXEL_GRID    0 = the k-NN estimate (default); 1 = the reference's
 *                      voxel-grid estimate (photonMappingKernel.cu:314-389,
 *                      :831-870): both maps splatted into 32^3 grids over the
 *                      room, a wall point gets the photon flux of the 6x6
 *                      window of its wall's boundary layer around its voxel
 *                      over the window's area, times albedo / pi (spheres get
 *                      none), a medium point the flux of the 6x6x6 interior
 *                      window over its volume, times 1 / (4 pi); 2 = as 1 with
 *                      the windows of the 4 nearest voxel centres blended
 *                      bilinearly (:442-583).  Needs pm_emit (or
 *                      pm_set_records) only; the grids are built by the first
 *                      pm_render / pm_render_crop after new records, which
 *                      pm_render_crop_noalloc needs (PM_ERR_STATE before). */
enum {
  PM_OPT_GATHER_MODE = 1,
  PM_OPT_BLOCKS_PER_CU = 2,
  PM_OPT_CACHE_FUSED_INFER = 3,
  PM_OPT_CACHE_INFER_CHUNK = 4,
  PM_OPT_TILE_REGIONS = 5,
  PM_OPT_SAMPLE_LANES = 6,
  PM_OPT_RENDER_MEDIA = 7,
  PM_OPT_VOXEL_GRID = 8,
  PM_OPT_TOP_LEVELS = 9,
  PM_OPT_CACHE_NEAR_GROUPS = 10,
  PM_OPT_CACHE_EMA = 11,
  PM_OPT_CACHE_SEEDS = 12
};
int pm_set_option(pm_context* ctx, int option, int value);

/* ------------------------------------------------------------------ learned radiance cache
 * A small MLP (9 inputs -> 12-band sin/cos encoding -> 4 x 128 ReLU -> RGB,
 * predicting log(1 + radiance)) trained online on the kNN gather's own
 * estimates.  BF16 weights on MFMA with FP32 accumulation; FP32 master
 * weights and Adam state.  Inputs are rows of 9 floats: position, direction,
 * normal (zero for volume queries).  No reference counterpart. */
uint32_t pm_cache_param_count(void);
/* (Re)initialise the weights deterministically from `seed`; clears Adam state
 * and restarts the cache render's frame count (which seeds the draws of its
 * training subsets), so that a reset context repeats the same frames. */
int pm_cache_reset(pm_context* ctx, uint32_t seed);
/* Copy the FP32 parameters to (set = 0) or from (set = 1) host memory; the
 * flat layout is, per layer l = 0..4, W_l [nout x kin] row-major then b_l,
 * with (kin, nout) = (224, 128), (128, 128) x 3, (128, 16); padding zero. */
int pm_cache_params(pm_context* ctx, float* params_host, uint32_t count, int set);
/* Loss gradient (same layout as the parameters) of mean((y - log(1+t))^2)
 * over n <= 2^20 device rows, without updating; loss_out may be NULL. */
int pm_cache_grad(pm_context* ctx, const float* inputs_dev, const float* targets_dev, uint32_t n,
                  float* grad_host, uint32_t count, double* loss_out);
/* Adam steps (one per `batch` rows, lr) over n device rows; mean loss out. */
int pm_cache_train(pm_context* ctx, const float* inputs_dev, const float* targets_dev, uint32_t n,
                   uint32_t batch, float lr, double* loss_out);
/* Radiance predictions (n x 3 FP32) for n device input rows. */
int pm_cache_predict(pm_context* ctx, const float* inputs_dev, uint32_t n, float* rgb_dev);

/* Cache render settings: fallback threshold on the mean (over a 16^3 cell's
 * training rows of this frame, per query kind, per colour channel) squared
 * error of the predicted log(1 + radiance), training samples per frame,
 * Adam batch and learning rate, march points per cached volume query (one
 * query, placed at the weight-weighted centroid of up to `march_stride`
 * consecutive points of a ray segment, stands for all of them with their
 * summed path weights; 1 = every point; within 4 x march_stride steps of the
 * light every point is its own query and is gathered exactly), and whether
 * surface queries use the cache (0: always gathered exactly).
 * Defaults 0.05, 2^18, 2^14, 1e-3, 1, 0. */
int pm_cache_configure(pm_context* ctx, float threshold, uint32_t train_samples, uint32_t batch, float lr,
                       int march_stride, int cache_surface);

typedef struct pm_cache_stats {
  uint64_t records;        /* radiance queries of the frame */
  uint64_t fallbacks;      /* of which gathered exactly (kNN) */
  double train_loss;       /* mean loss of this frame's training steps */
  /* phase times on the frame's stream; when surface records are not cached,
   * their exact gather runs on a second stream from the start of training,
   * so it overlaps ms_train and ms_infer and ms_fallback holds the rest */
  double ms_record, ms_train, ms_infer, ms_fallback, ms_accumulate, ms_total;
  uint32_t frame;          /* frames rendered with this cache */
  uint32_t inferred_volume; /* of the inferred rows, volume records on the 5-step layer 1 */
  uint64_t inferred;       /* rows through the network (the volume records when surface records are not cached) */
} pm_cache_stats;

/* Render the crop (x0, y0, w, h) of the configured frame through the cache:
 * record all radiance queries, train on train_samples of them (exact kNN
 * targets), predict the rest, re-gather exactly where the cell error is above
 * the threshold, accumulate.  Same outputs as pm_render_crop.  While the
 * records and the photon positions stay the same from one call to the next,
 * the exact re-gathers start from each record's k-th-neighbour bucket of the
 * previous call (PM_OPT_CACHE_SEEDS; the same values, fewer passes). */
int pm_render_cache_crop(pm_context* ctx, int x0, int y0, int w, int h, float* rgb_dev, pm_uchar4* out_dev);
int pm_cache_stats_get(pm_context* ctx, pm_cache_stats* out);

/* ------------------------------------------------------------------ batched scenes
 * Many independent scenes of uneven photon counts and resolutions rendered
 * per frame by one hipGraph launch: batched emission over all scenes (one
 * count pass, one scan, one write pass), one kd-tree forest build over every
 * photon map, and one persistent render kernel fed by a device work queue of
 * (scene, 8x8 tile) items.  Every scene's records, trees and pixels are
 * bit-identical to its own pm_create/pm_emit/pm_build/pm_render run with the
 * same leaf size.  The workgroup size and heap slots never change a result;
 * the leaf size changes only the order in which the (identical) neighbours'
 * contributions are summed.  No reference counterpart (the reference renders
 * one scene). */
typedef struct pm_batch pm_batch;

typedef struct pm_batch_tuning {
  int leaf_size;      /* kd-tree leaf size L (1..64) of every map */
  int block_threads;  /* render workgroup size: 64, 128, 192 or 256 */
  int heap_slots;     /* LDS heap keys per lane: 32 or 64 (>= every scene's k) */
  int heap_key_bytes; /* heap key width: 8 (d^2, index), 4 (d^2) or 2 (d^2 bucket); see PM_OPT_GATHER_MODE */
  int gather_mode;    /* 0: the heap gather of heap_key_bytes (modes 1, 2, 4); 7: gather mode 7, which needs
                         heap_key_bytes 2 and heap_slots 64 (its bucket counters take that LDS) */
  int top_levels;     /* LDS tile, gather mode 7 only: the top 0..7 levels of the current scene's two kd-trees
                         (2^levels - 1 nodes each, 48 B per node) staged in each wave's LDS, 0 = none */
} pm_batch_tuning;

typedef struct pm_batch_stats {
  uint32_t n_scenes;
  uint32_t work_items;                       /* 8x8 tiles of all scenes */
  uint64_t photons, surface_records, volume_records, pixels;
  uint64_t surface_queries, volume_queries;  /* radiance queries of the last frame */
  double ms_frame;                           /* last pm_batch_submit: graph launch to completion */
  double ms_setup;                           /* sizing pass, tables and graph capture */
  pm_batch_tuning tuning;
  uint32_t graph_nodes;
  uint32_t pad0;
} pm_batch_stats;

/* Scene `index` of the batched workload generated from `seed`: the unit room
 * with 2-256 random spheres (60% Lambertian / 25% mirror / 15% dielectric,
 * IOR 1.5), 10^4..2*10^6 photons (log-uniform), image side 128..1024,
 * homogeneous medium sigma_s U[0, 0.3]; k = 32, r <= 0.05, 1 spp. */
int pm_batch_scene_config(uint32_t seed, int index, pm_config* out);
/* n_scenes <= 128 configurations with n_emit > 0 (photon-count mode) and the
 * same storage, estimator, transmittance precision and media switch; k <= 64.
 * Runs the sizing pass and captures the frame graph. */
int pm_batch_create(const pm_config* cfgs, int n_scenes, pm_batch** out);
int pm_batch_destroy(pm_batch* b);
int pm_batch_set_tuning(pm_batch* b, const pm_batch_tuning* t);
/* Times the kd-build and a 1-in-`sample_stride` subset of the render work
 * items for a staged sweep: heap key width {8, 4, 2} x heap slots {32 (when
 * every k <= 32), 64} and gather mode 7, then leaf size {4, 8, 16, 32}, then
 * workgroup size {64, 128, 256}, then (when gather mode 7 won) the LDS tile
 * top_levels {0, 3, 5, 7}, each stage from the best so far; keeps the
 * fastest and returns it (and its time). */
int pm_batch_autotune(pm_batch* b, int sample_stride, pm_batch_tuning* best_out, double* best_ms_out);
/* One frame of every scene (one graph launch); blocking. */
int pm_batch_submit(pm_batch* b);
/* The same frame with its kernels launched one by one on the batch's stream
 * instead of through the graph, timed by events between its phases: the
 * emission (trace, scans, tally), the forest build (with node boxes and the
 * tree-ordered payload) and the render.  Blocking; leaves the same images,
 * records and trees as pm_batch_submit, and no statistics.  No counterpart in
 * the reference (its single scene is timed by simpleGLMain.cpp:34-49). */
int pm_batch_time_phases(pm_batch* b, double* ms_emit, double* ms_build, double* ms_render);
/* Scene `scene`'s image of the last frame: rgb_dev w*h*3 FP32 and/or out_dev
 * w*h uchar4 (either may be NULL). */
int pm_batch_image(pm_batch* b, int scene, float* rgb_dev, pm_uchar4* out_dev);
/* Per-scene record counts and flux tally of the last frame (timing fields 0). */
int pm_batch_scene_stats(pm_batch* b, int scene, pm_stats* out);
int pm_batch_stats_get(pm_batch* b, pm_batch_stats* out);
int pm_batch_get_records(pm_batch* b, int scene, int map, float* pos_host, float* power_host, float* dir_host,
                         uint32_t n);
int pm_batch_get_tree(pm_batch* b, int scene, int map, uint32_t* perm_host, float* split_host, uint32_t* axis_host,
                      uint32_t n, uint32_t n_internal);

/* Inspection (copies to host memory; for tests and tools). */
int pm_get_records(pm_context* ctx, int map, float* pos_host, float* power_host, float* dir_host, uint32_t n);
/* Replaces map `map`'s records with n caller records (host arrays of 3 floats
 * per record; power and direction rounded to BF16 under BF16 storage), whose
 * flux normalisation is light_power / n_emit; pm_build must follow.  For
 * measurements on given point sets (the other map is kept).  No reference
 * counterpart (the reference has no photon map API). */
int pm_set_records(pm_context* ctx, int map, const float* pos_host, const float* power_host, const float* dir_host,
                   uint32_t n, uint32_t n_emit);
int pm_get_tree(pm_context* ctx, int map, uint32_t* perm_host, float* split_host, uint32_t* axis_host,
                uint32_t n, uint32_t n_internal);
/* The voxel-grid estimate's summed-volume table of map `map` (PM_OPT_VOXEL_GRID):
 * n = 3 * 33^3 entries, entry 3 * ((z * 33 + y) * 33 + x) + channel = the
 * record power (in units of 2^-24) of the voxels [0, x) x [0, y) x [0, z),
 * modulo 2^64.  PM_ERR_STATE until a voxel-grid render has built it. */
int pm_get_voxel_table(pm_context* ctx, int map, uint64_t* table_host, size_t n);
/* kNN over a map's kd-tree for nq device query points (xyz, FP32): idx_dev and
 * d2_dev receive nq*k entries sorted by (d2, photon index), padded with
 * 0xffffffff / +inf; count_dev (may be NULL) receives the number found. */
int pm_knn(pm_context* ctx, int map, const float* queries_dev, uint32_t nq, int k, float rcap,
           uint32_t* idx_dev, float* d2_dev, uint32_t* count_dev);

/* Device memory helpers for hosts without a HIP runtime binding (ctypes). */
int pm_device_alloc(void** ptr, size_t bytes);
int pm_device_free(void* ptr);
int pm_memcpy_htod(void* dst_dev, const void* src_host, size_t bytes);
int pm_memcpy_dtoh(void* dst_host, const void* src_dev, size_t bytes);
int pm_device_count(int* n);
/* Plain device-to-device copy kernel (16-byte vector loads/stores) on the
 * library's stream; reads and writes exactly `bytes` (a multiple of 16).  Used
 * to calibrate hardware memory counters against a known byte count. */
int pm_copy_kernel(void* dst_dev, const void* src_dev, size_t bytes);
const char* pm_status_string(int status);

#ifdef __cplusplus
}
#endif

#endif /* PHOTON_MAPPER_H */
