// pm_kernels.h -- host-side launchers of the gfx950 kernels and the device
// data layout they share.  Internal to libpm.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../core/pm_kdtree.h"
#include "../core/pm_transport.h"
#include "../core/pm_voxel.h"

namespace pm {

// Photon map on the device, structure of arrays.  Records are in (photon id,
// event) order; `payload` holds power and travel direction per record:
//   FP32 storage: 8 floats  (pr, pg, pb, dx | dy, dz, 0, 0)   = 32 B
//   BF16 storage: 8 bf16    (pr, pg, pb, dx, dy, dz, 0, 0)    = 16 B
// The kd-tree keeps the positions again in tree order as float4
// (x, y, z, bits(record index)) so a leaf is one contiguous, coalesced read.
struct DevMap {
  uint32_t n;        // records
  uint32_t cap;      // allocated records
  uint32_t n_emit;   // photons emitted for this map's normalisation
  int bf16;
  float* x;
  float* y;
  float* z;
  void* payload;
  // kd-tree
  uint32_t L, nl;
  float4* tpts;      // n, tree order
  KdNode* nodes;     // nl - 1
  uint2* nrange;     // nl - 1: point range [lo, hi) of each internal node
  float4* bmin;      // 2 nl - 1 (internal nodes, then leaves): tight box of each node's points, or null
  float4* bmax;
  void* tpayload;    // payload in tree order (tpayload[i] = payload[record of tpts[i]]), or null
  bool built;
};

// Tree-order point arrays carry this many entries past the last point, so that
// a leaf batch reads up to 8 consecutive entries from one address without
// clamping; entries past a leaf's end are read but never used.
static constexpr uint32_t kTptsPad = 8;

// Gather parameters of one map (passed by value to kernels).
struct GatherMap {
  const float4* tpts;
  const KdNode* nodes;
  const uint2* nrange;
  const float4* bmin;  // tight node boxes (2 nl - 1), or null
  const float4* bmax;
  const void* payload;
  const void* tpayload;  // payload in tree order, or null
  uint32_t n, nl;
  int bf16;
  int k;
  float rcap2;
  float scale[3];  // light power / n_emit
  // top-of-tree tile: the first top_count internal nodes' records, staged in
  // LDS by the gather kernels that take one (0: none, every read from HBM)
  const KdNode* top_nodes = nullptr;
  const uint2* top_nrange = nullptr;
  const float4* top_bmin = nullptr;
  const float4* top_bmax = nullptr;
  uint32_t top_count = 0;
};
// LDS bytes of a top-of-tree tile of `levels` levels (2^levels - 1 nodes):
// boxes (2 x 16 B), node record (8 B) and point range (8 B) per node.
PM_HD size_t top_tile_bytes(int levels) { return levels > 0 ? 48u * (((size_t)1 << levels) - 1u) : 0u; }
constexpr int kMaxTopLevels = 7;

// scan / sort (pm_scan_sort.hip)
size_t scan_temp_words(size_t n);
hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, size_t n, uint32_t* temp, uint32_t* total_dev,
                          hipStream_t st);
size_t radix_temp_words(size_t n);
hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, size_t n, uint32_t* temp, hipStream_t st);
// One stable pass on the 8-bit digit at `shift`: (kin, vin) -> (kout, vout).
hipError_t radix_pass(const uint32_t* kin, const uint32_t* vin, size_t n, int shift, uint32_t* temp, uint32_t* kout,
                      uint32_t* vout, hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t st);
hipError_t launch_copy16(const void* src, void* dst, size_t bytes, hipStream_t st);

// trace (pm_trace.hip)
hipError_t launch_trace_count(const Scene& sc, const TraceParams& tp, uint32_t pid0, uint32_t npid,
                              uint32_t* cnt_s, uint32_t* cnt_v, hipStream_t st);
hipError_t launch_find_cap(const uint32_t* off, const uint32_t* cnt, uint32_t n, uint32_t cap, uint32_t* out,
                           hipStream_t st);
hipError_t launch_trace_write(const Scene& sc, const TraceParams& tp, uint32_t npid, const uint32_t* off_s,
                              const uint32_t* off_v, DevMap ms, DevMap mv, double* tally_partial, hipStream_t st);
hipError_t launch_tally_final(const double* partial, uint32_t nblocks, double* out9, hipStream_t st);
uint32_t trace_write_blocks(uint32_t npid);

// kd-tree (pm_kdbuild.hip)
static constexpr uint32_t KD_LEAF = 0xffffffffu;
// A forest of canonical trees over one concatenated point array.  Tables are
// device arrays indexed by global node id; null optional tables mean a single
// tree in heap order (children 2h+1, 2h+2 below nnodes, point base 0).
struct KdForest {
  uint32_t n;                   // points of all trees
  const float* x;
  const float* y;
  const float* z;
  float4* tpts;                 // [n] (x, y, z, bits(tree-local point index)) in tree order
  KdNode* nodes;                // [nnodes]
  uint2* nrange;                // [nnodes] tree-local point range [lo, hi)
  uint32_t ntrees;              // 1 .. 256
  uint32_t nnodes;              // internal nodes of all trees
  const uint32_t* point_tree;   // [n] tree of each point (null: one tree)
  const uint32_t* tree_pbase;   // [ntrees] first point of each tree
  const uint32_t* tree_root;    // [ntrees] root node id, KD_LEAF without internal nodes
  const uint32_t* node_lo;      // [nnodes] global point range [lo, hi)
  const uint32_t* node_hi;
  const uint32_t* node_lc;      // left point count
  const uint32_t* node_lb;      // left counts of the same level's earlier nodes (all trees)
  const uint32_t* node_child;   // [2 nnodes] child ids or KD_LEAF (optional)
  const uint32_t* node_pbase;   // [nnodes] first point of the node's tree (optional)
  const uint32_t* level_nodes;  // node ids in level order (optional: identity)
};
size_t kdforest_temp_words(uint32_t n, uint32_t nnodes);
// level_off: host array of nlevels + 1 offsets into level_nodes.
hipError_t kd_build_forest(const KdForest& F, const uint32_t* level_off, uint32_t nlevels, uint32_t* temp,
                           hipStream_t st);
size_t kdbuild_temp_words(uint32_t n, uint32_t L);
// Shape tables of the ni = nl-1 internal nodes, in device memory: point range
// [lo, hi), left count lc, and lb = left counts of the same level's earlier
// nodes, each at a stride of ni + 1 words (kd_shape_table_words in all).  They
// depend on n and L only, so a caller keeps them across builds of that size.
size_t kd_shape_table_words(uint32_t n, uint32_t L);
hipError_t kd_build_impl(DevMap& m, uint32_t* temp, const uint32_t* tables, hipStream_t st);
// Tight boxes of every node's points into m.bmin / m.bmax (when allocated).
hipError_t kd_node_boxes(const DevMap& m, hipStream_t st);
// The payload in tree order into m.tpayload (when allocated).
hipError_t kd_tree_payload(const DevMap& m, hipStream_t st);
// The same two for a forest (trees laid out one after another).
struct KdTreeDesc {
  uint32_t n, nl;  // points, leaves
  uint32_t nbase;  // first internal node in the forest's node arrays
  uint32_t pbase;  // first point
  uint32_t bbase;  // first box: a tree has 2 nl - 1 boxes, heap nodes 0 .. 2 nl - 2
  uint32_t pad;
};
struct KdForestBoxes {
  const float4* tpts;
  const KdNode* nodes;
  const uint2* nrange;           // tree-local point ranges
  const KdTreeDesc* trees;
  uint32_t ntrees;
  const uint32_t* leaf_pref;     // [ntrees + 1] leaves of the earlier trees
  const uint32_t* level_pref;    // [depth][ntrees + 1] internal nodes at level d of the earlier trees
  uint32_t depth;                // internal levels of the deepest tree
  uint32_t nleaves;
  float4* bmin;
  float4* bmax;
};
// level_total (host): [depth] internal nodes at level d of all trees
hipError_t kd_forest_boxes(const KdForestBoxes& f, const uint32_t* level_total, hipStream_t st);
// tpayload[i] = payload[tree_pbase[point_tree[i]] + tree-local index of tpts[i]]
hipError_t kd_forest_payload(const float4* tpts, const uint32_t* point_tree, const uint32_t* tree_pbase,
                             const void* payload, void* tpayload, uint32_t n, int bf16, hipStream_t st);

// gather (pm_gather.hip)
hipError_t launch_knn(GatherMap g, const float* q, uint32_t nq, uint32_t* idx, float* d2, uint32_t* count,
                      hipStream_t st);
struct RenderLaunch {
  Scene sc;
  RenderParams rp;
  GatherMap gs, gv;
  int estimator;  // 0 box, 1 cone
  float exposure;
  int x0, y0, w, h;
  float* rgb;      // w*h*3 or null
  uchar4* u8;      // w*h or null
  unsigned long long* counters;  // [surface queries, volume queries]
  int blocks_per_cu;             // persistent grid: workgroups per compute unit
  int gather_mode;               // 0: wave per query; lane per query: 1 8-byte, 2 4-byte, 3 2-byte heap
  int tile_regions;              // lane modes: 0 raster tile order; rx = 1, 2, 4, 8: rx x (8 / rx) per-XCD regions
  int sample_lanes;              // mode 7: a pixel's spp samples in adjacent lanes (spp 2, 4, 8 or 16)
  int top_levels;                // mode 7: levels of each tree staged in LDS per wave (0: none)
};
hipError_t launch_render(const RenderLaunch& r, hipStream_t st);
hipError_t launch_render_lane(const RenderLaunch& r, hipStream_t st);
hipError_t launch_render_lane2(const RenderLaunch& r, hipStream_t st);
hipError_t launch_render_lane4(const RenderLaunch& r, hipStream_t st);
hipError_t launch_render_lane7(const RenderLaunch& r, hipStream_t st);

// voxel-grid compatibility mode (pm_voxel.hip): both maps splatted into
// [32^3][3] fixed-point grids, then their [33^3][3] summed-volume tables
hipError_t launch_voxel_grids(const DevMap& ms, const DevMap& mv, unsigned long long* grid_s,
                              unsigned long long* grid_v, unsigned long long* sat_s, unsigned long long* sat_v,
                              hipStream_t st);
hipError_t launch_render_voxel(const RenderLaunch& r, const VoxelGrids& G, hipStream_t st);

// batched scenes (pm_trace.hip, pm_gather_lane.hip)
// Emission: scene s owns the trace blocks [block0, block0 + nblocks), whose
// lanes trace its photon ids 0 .. n_emit-1 (the one-scene block partition).
struct BatchTraceScene {
  Scene sc;
  TraceParams tp;
  uint32_t n_emit;
  uint32_t block0, nblocks;
  uint32_t pad;
};
hipError_t launch_trace_count_batch(const BatchTraceScene* scenes, const uint32_t* block_scene, uint32_t nblocks,
                                    uint32_t* cnt_s, uint32_t* cnt_v, hipStream_t st);
hipError_t launch_trace_write_batch(const BatchTraceScene* scenes, const uint32_t* block_scene, uint32_t nblocks,
                                    const uint32_t* off_s, const uint32_t* off_v, DevMap ms, DevMap mv,
                                    double* tally_partial, hipStream_t st);
// out9[9 s + c]: scene s's tally, summed over its blocks in the one-scene order.
hipError_t launch_tally_final_batch(const double* partial, const BatchTraceScene* scenes, uint32_t nscenes,
                                    double* out9, hipStream_t st);
uint32_t trace_threads();
struct BatchRenderScene {
  Scene sc;
  RenderParams rp;
  GatherMap gs, gv;
  float exposure;
  int w, h;
  int pad;
  uint64_t img_off;  // first pixel of the scene's image in the batch image arena
};
struct BatchRenderLaunch {
  const BatchRenderScene* scenes;
  const uint2* items;              // work items (scene, 8x8 tile), most expensive first
  uint32_t nitems;
  uint32_t* head;                  // work-queue counter, zero at launch
  float* rgb;                      // image arena, 3 floats per pixel
  unsigned long long* counters;    // [surface queries, volume queries]
  int heap_slots;                  // LDS heap keys per lane (>= every scene's k)
  int key_bytes;                   // heap key width: 8, 4 or 2 (gather modes 1, 2, 4)
  int counting;                    // 1: gather mode 7 (key_bytes 2, heap_slots 64; maps with boxes and tpayload)
  int top_levels;                  // gather mode 7: levels of each scene's trees staged in LDS per wave
};
size_t render_batch_lds_bytes(int block_threads, int heap_slots, int key_bytes, int top_levels);
int render_batch_occupancy(int fp64, int cone, int block_threads, int heap_slots, int key_bytes, int counting,
                           int top_levels);  // per CU
hipError_t launch_tonemap(const float* rgb, size_t npix, float exposure, uchar4* out, hipStream_t st);
hipError_t launch_gather_u32(const uint32_t* src, const uint32_t* idx, uint32_t n, uint32_t* dst, hipStream_t st);
hipError_t launch_render_batch(const BatchRenderLaunch& r, int fp64, int cone, int block_threads, int grid,
                               hipStream_t st);

// learned-cache render (pm_cache_render.hip)
enum : int { CACHE_GRID = 16 };  // error-tracker cells per axis (x 2 kinds)
// Record order of a crop's pixels: 4x4 blocks in row-major block order,
// row-major inside a block, so that consecutive records (and the rows of a
// wave in the record lists) come from a compact patch of the image.  The
// per-pixel counts and offsets are indexed by this order over the crop padded
// to whole blocks (cache_order_size slots; padding slots hold no records).
PM_HD uint32_t cache_pixel_order(int x, int y, int w) {
  const uint32_t bx = (uint32_t)((w + 3) / 4);
  return ((uint32_t)(y >> 2) * bx + (uint32_t)(x >> 2)) * 16u + (uint32_t)((y & 3) * 4 + (x & 3));
}
PM_HD size_t cache_order_size(int w, int h) { return (size_t)((w + 3) / 4) * (size_t)((h + 3) / 4) * 16u; }
// The error-tracker cell of a record at (x, y, z): its 16^3 cell (clamped to
// the grid), plus 4096 for volume records.
PM_HD uint32_t err_cell_of(float x, float y, float z, bool surf) {
  const float p[3] = {x, y, z};
  int c[3];
  for (int a = 0; a < 3; ++a) {
    float v = p[a] * (float)CACHE_GRID;
    int iv = (int)v;
    c[a] = iv < 0 ? 0 : (iv >= CACHE_GRID ? CACHE_GRID - 1 : iv);
  }
  return (uint32_t)((((surf ? 0 : 1) * CACHE_GRID + c[2]) * CACHE_GRID + c[1]) * CACHE_GRID + c[0]);
}
// The cell of a query row (position, direction, normal or 0).
PM_HD uint32_t err_cell(const float* p) {
  return err_cell_of(p[0], p[1], p[2], !(p[6] == 0.0f && p[7] == 0.0f && p[8] == 0.0f));
}
struct CacheRecordLaunch {
  Scene sc;
  RenderParams rp;
  int x0, y0, w, h;
  int write;                 // 0: count queries per pixel, 1: write records
  uint32_t* counts;          // [cache_order_size(w, h)], by cache_pixel_order
  const uint32_t* offsets;   // [cache_order_size(w, h)]
  float* in;                 // [N x 9]
  float* wt;                 // [N x 3]
  uint16_t* code;            // [N] record codes (error cell, kind, near-light bit; record_code)
  int stride;                // march points per volume record
  float near2;               // squared near-light radius: no merging inside
};
enum : int { CACHE_CELLS = 2 * CACHE_GRID * CACHE_GRID * CACHE_GRID };
// The two record lists of a cache frame, each in record order, built from the
// record codes: list 0 then list 1 (record_class in pm_cache_render.hip).
struct CacheListRule {
  int fallback;              // 0: list 0 = surface records, list 1 = volume records;
                             // 1: the exact fallbacks, surface ones in list 0, volume ones in list 1
  int cache_surface;         // 0: every surface record falls back
  const uint8_t* cell_fb;    // fallback lists: per error cell, 1 when its records fall back
};
struct CacheValueLaunch {
  GatherMap gs, gv;
  float hg_g;
  const float* in;           // query rows (9 floats)
  const uint32_t* list;      // rows to evaluate (null: 0..n-1)
  uint32_t n;
  float* out;                // values (3 floats)
  int out_by_list;           // 1: out[list[i]], 0: out[i]
  int mode7;                 // 1: the gather of mode 7 (k_values7), 0: the 2-byte heap gather (k_values)
  int kind;                  // rows are 0: of either kind, 1: all surface queries, 2: all volume queries
  uint16_t* seed_q;          // k_values7, rows with out_by_list: per record, the k-th neighbour's bucket
                             // of its last gather (read as the counting pass's seed, then rewritten;
                             // 0xffff: none); null: heap passes
};
hipError_t launch_cache_record(const CacheRecordLaunch& r, hipStream_t st);
hipError_t launch_cache_values(const CacheValueLaunch& v, hipStream_t st);
hipError_t launch_cache_select(const float* in, const uint32_t* list_s, uint32_t n_s, const uint32_t* list_v,
                               uint32_t n_v, uint32_t count, uint32_t seed, uint32_t frame, float* out, hipStream_t st);
// *acc += v[0] + ... + v[n - 1] in double, in a fixed order; `part` holds
// sum_rows_parts() doubles of scratch
size_t sum_rows_parts();
hipError_t launch_sum_rows(const float* v, uint32_t n, double* acc, double* part, hipStream_t st);
// *out = a 64-bit fingerprint of the positions of n points (equal arrays,
// equal fingerprints)
hipError_t launch_point_hash(const float4* p, uint32_t n, unsigned long long* out, hipStream_t st);
// *total = cnt[0] + ... + cnt[n - 1] in 64 bits
hipError_t launch_count_total(const uint32_t* cnt, uint32_t n, unsigned long long* total, hipStream_t st);
// cell_fb[c] = 1 when error cell c saw no training row or its mean squared
// log error exceeds the threshold (CACHE_CELLS bytes).
hipError_t launch_cache_cell_fb(const unsigned long long* esum, const unsigned int* ecnt, float threshold,
                                uint8_t* cell_fb, hipStream_t st);
// List building over n record codes in cache_list_blocks(n) blocks: counts
// [2 x blocks] (list 0's per block, then list 1's); after an exclusive scan of
// the counts into offs, the scatter writes list 0 at [0, total_0) and list 1
// at [total_0, total_0 + total_1) of `list`.
size_t cache_list_blocks(uint64_t n);
// code[i] = the record code of query row i (its error cell and kind; no near-light bit)
hipError_t launch_cache_row_codes(const float* in, uint64_t n, uint16_t* code, hipStream_t st);
hipError_t launch_cache_list_count(const uint16_t* code, uint64_t n, const CacheListRule& rule, uint32_t* counts,
                                   hipStream_t st);
hipError_t launch_cache_list_scatter(const uint16_t* code, uint64_t n, const CacheListRule& rule,
                                     const uint32_t* offs, uint32_t* list, hipStream_t st);
// Per crop pixel (row-major, w x h): the records [off[o], off[o] + cnt[o]) of
// its record-order slot o = cache_pixel_order(x, y, w).
hipError_t launch_cache_accumulate(const uint32_t* off, const uint32_t* cnt, int w, int h, const float* wt,
                                   const float* val, int spp, float exposure, float* rgb, uchar4* u8, hipStream_t st);

}  // namespace pm
