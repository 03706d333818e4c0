// pm_kdbuild.hip -- GPU construction of the canonical left-balanced kd-tree
// (shape and split rule in core/pm_kdtree.h).  The reference has no spatial
// structure (a fixed 32^3 voxel grid, photonMappingKernel.cu:15-17,260-267).
//
// Method (O(n log n), every step data-parallel and deterministic):
//  1. For each axis, radix-sort (key(coord), photon index) pairs: three index
//     lists, each ordered by (coordinate, index).
//  2. Level by level, for every internal node of the level at once:
//     bounding box = first/last entries of the three lists inside the node's
//     range; split axis = largest extent; median = the list entry at the
//     shape-determined split position.  Then each list is stably partitioned
//     inside every node (flag = "orders before the median"), which keeps all
//     three lists sorted per child.  Stable partition = global exclusive scan
//     of the flags; the flags before a node's start are known from the shape
//     (the left counts of the level's earlier nodes), so no segmented scan is
//     needed.
//  3. After the last level the x-list gives the leaf order (x key, index);
//     positions are gathered into float4 (x, y, z, index) in tree order.
// The points are first renumbered in Morton order (a 30-bit Morton code per
// point, radix-sorted; within each tree of a forest), and the build works on
// the renumbered ids and a copy of the coordinates in that order: the points
// of a node are then close in memory as well as in space, so the per-level
// coordinate gathers of step 2 hit cached lines.  The photon index stays the
// tie-breaker of every order (the lists are sorted by (key, photon index)
// before being renumbered), so the tree is the same.
#include <string.h>

#include <algorithm>
#include <vector>

#include "pm_kernels.h"
#include "pm_wave.h"

namespace pm {

static constexpr int KB_THREADS = 256;
static constexpr int KB_ITEMS = 4;
static constexpr int KB_TILE = KB_THREADS * KB_ITEMS;

__device__ __forceinline__ float coord_of(const float* x, const float* y, const float* z, uint32_t i, int a) {
  return a == 0 ? x[i] : (a == 1 ? y[i] : z[i]);
}

// 10 bits spread to every third bit of 30 (bit b -> bit 3b).
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__device__ __forceinline__ uint32_t quant10(float c) {
  c = fminf(fmaxf(c, 0.0f), 1.0f);  // the room is the unit cube; NaN -> 0
  const uint32_t q = (uint32_t)(c * 1024.0f);
  return q > 1023u ? 1023u : q;
}

// Morton code of each point over the unit cube, with its index.
__global__ void k_morton_keys(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                              uint32_t n, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = spread3(quant10(x[i])) | (spread3(quant10(y[i])) << 1) | (spread3(quant10(z[i])) << 2);
  vals[i] = i;
}

// The renumbering: inv[orig[j]] = j and the coordinates in the new order.
__global__ void k_renumber(const uint32_t* __restrict__ orig, uint32_t n, const float* __restrict__ x,
                           const float* __restrict__ y, const float* __restrict__ z, uint32_t* __restrict__ inv,
                           float* __restrict__ xr, float* __restrict__ yr, float* __restrict__ zr) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t i = orig[j];
  inv[i] = j;
  xr[j] = x[i];
  yr[j] = y[i];
  zr[j] = z[i];
}

// list[p] = inv[list[p]]: a list of photon indices as renumbered ids.
__global__ void k_map_ids(uint32_t* __restrict__ list, uint32_t n, const uint32_t* __restrict__ inv) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) list[p] = inv[list[p]];
}

__global__ void k_make_keys(const float* __restrict__ c, uint32_t n, uint32_t* __restrict__ keys,
                            uint32_t* __restrict__ vals) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = float_key(c[i]);
  vals[i] = i;
}

// A forest of canonical trees over one concatenated point array (one tree
// for the single-map build; one per photon map of every scene in a batch).
// Node ids are global; seg[p] holds the node id of point position p at the
// current level, or KD_LEAF once its subtree is a leaf.
__device__ __forceinline__ uint32_t kd_child(const KdForest& F, uint32_t h, bool left) {
  if (F.node_child) return F.node_child[2u * h + (left ? 0u : 1u)];
  uint32_t c = 2u * h + (left ? 1u : 2u);
  return c < F.nnodes ? c : KD_LEAF;
}

// One thread per internal node of the level: level_nodes[first .. first+count)
// (identity ids when level_nodes is null).
__global__ void k_level_nodes(KdForest F, const uint32_t* __restrict__ l0, const uint32_t* __restrict__ l1,
                              const uint32_t* __restrict__ l2, uint32_t first, uint32_t count,
                              uint32_t* __restrict__ node_med, uint32_t* __restrict__ node_axis) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  uint32_t h = F.level_nodes ? F.level_nodes[first + t] : first + t;
  uint32_t lo = F.node_lo[h], hi = F.node_hi[h], lc = F.node_lc[h];
  const uint32_t* L[3] = {l0, l1, l2};
  float ext[3];
  for (int a = 0; a < 3; ++a) {
    float mn = coord_of(F.x, F.y, F.z, L[a][lo], a);
    float mx = coord_of(F.x, F.y, F.z, L[a][hi - 1], a);
    ext[a] = mx - mn;
  }
  int axis = 0;
  float best = ext[0];
  if (ext[1] > best) { best = ext[1]; axis = 1; }
  if (ext[2] > best) { best = ext[2]; axis = 2; }
  uint32_t med = L[axis][lo + lc];
  KdNode nd;
  nd.split = coord_of(F.x, F.y, F.z, med, axis);
  nd.meta = (uint32_t)axis | (lc << 2);
  F.nodes[h] = nd;
  const uint32_t pb = F.node_pbase ? F.node_pbase[h] : 0u;
  F.nrange[h] = make_uint2(lo - pb, hi - pb);
  node_med[h] = med;
  node_axis[h] = (uint32_t)axis;
}

struct LevelArgs {
  KdForest F;
  const uint32_t* lin[3];
  uint32_t* lout[3];
  const uint32_t* seg;
  uint32_t* seg_out;
  const uint32_t* node_med;
  const uint32_t* node_axis;
  uint32_t nblocks;
  uint32_t* blockcnt;  // [3][nblocks]
  uint8_t* flags;      // [n] per position: bit b = flag of list b, bit 7 = in a node of this level
  const uint32_t* orig;  // photon index of each renumbered id (the tie-breaker)
};

// Flag of list b at position p: does the element order before the node's
// median under (key on the node's axis, index)?  On the node's own axis the
// list is already split at the median, so the flag is "p < median position".
__device__ __forceinline__ bool part_flag(const LevelArgs& A, int b, uint32_t p, uint32_t h) {
  uint32_t a = A.node_axis[h];
  if ((uint32_t)b == a) return p < A.F.node_lo[h] + A.F.node_lc[h];
  uint32_t i = A.lin[b][p];
  uint32_t m = A.node_med[h];
  uint32_t ki = float_key(coord_of(A.F.x, A.F.y, A.F.z, i, (int)a));
  uint32_t km = float_key(coord_of(A.F.x, A.F.y, A.F.z, m, (int)a));
  return ki < km || (ki == km && A.orig[i] < A.orig[m]);
}

// Counts the flags per block and keeps them (A.flags) for k_level_scatter:
// the flags of the two lists off the node's axis gather a coordinate of a
// point in another list's order, the costly part of a level.
__global__ void __launch_bounds__(KB_THREADS) k_level_count(LevelArgs A) {
  __shared__ uint32_t lw[KB_THREADS / 64];
  uint32_t cnt[3] = {0u, 0u, 0u};
  size_t base = (size_t)blockIdx.x * KB_TILE + (size_t)threadIdx.x * KB_ITEMS;
  for (int r = 0; r < KB_ITEMS; ++r) {
    size_t p = base + r;
    if (p >= A.F.n) break;
    uint32_t h = A.seg[p];
    uint8_t m = 0;
    if (h != KD_LEAF) {
      m = 0x80;
      for (int b = 0; b < 3; ++b)
        if (part_flag(A, b, (uint32_t)p, h)) {
          m |= (uint8_t)(1u << b);
          cnt[b] += 1u;
        }
    }
    A.flags[p] = m;
  }
  for (int b = 0; b < 3; ++b) {
    // reduce via wave sums + LDS
    uint32_t v = cnt[b];
    const int lane = lane_id();
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += lane_read(v, lane ^ d);
    if (lane == 0) lw[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t s = 0;
      for (int w = 0; w < KB_THREADS / 64; ++w) s += lw[w];
      A.blockcnt[(size_t)b * A.nblocks + blockIdx.x] = s;
    }
    __syncthreads();
  }
}

__device__ __forceinline__ uint32_t kb_block_excl(uint32_t v, uint32_t* lw) {
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = lane_read(x, lane >= d ? lane - d : lane);
    if (lane >= d) x += y;
  }
  if (lane == 63) lw[wave] = x;
  __syncthreads();
  uint32_t off = 0;
  for (int w = 0; w < wave; ++w) off += lw[w];
  __syncthreads();
  return off + x - v;
}

// blockoff: exclusive scan of blockcnt over the concatenated [3][nblocks]
// array; list b's own offsets are blockoff[b*nb + blk] - blockoff[b*nb].
// The flags of the level's earlier nodes number exactly their left counts
// (node_lb), so the left part of position p's node starts at lo + (run - lb).
__global__ void __launch_bounds__(KB_THREADS) k_level_scatter(LevelArgs A, const uint32_t* __restrict__ blockoff) {
  __shared__ uint32_t lw[KB_THREADS / 64];
  size_t base = (size_t)blockIdx.x * KB_TILE + (size_t)threadIdx.x * KB_ITEMS;
  uint32_t hs[KB_ITEMS];
  uint8_t fl[KB_ITEMS];  // bit b = flag of list b; bit 7 = position is in a node of this level
  uint32_t cnt[3] = {0u, 0u, 0u};
  for (int r = 0; r < KB_ITEMS; ++r) {
    size_t p = base + r;
    fl[r] = 0;
    hs[r] = KD_LEAF;
    if (p >= A.F.n) continue;
    hs[r] = A.seg[p];
    const uint8_t m = A.flags[p];  // from k_level_count
    for (int b = 0; b < 3; ++b) cnt[b] += (m >> b) & 1u;
    fl[r] = m;
  }
  uint32_t run[3];
  for (int b = 0; b < 3; ++b) {
    uint32_t boff = blockoff[(size_t)b * A.nblocks + blockIdx.x] - blockoff[(size_t)b * A.nblocks];
    run[b] = boff + kb_block_excl(cnt[b], lw);
  }
  for (int r = 0; r < KB_ITEMS; ++r) {
    size_t p = base + r;
    if (p >= A.F.n) break;
    uint32_t h = hs[r];
    if (!(fl[r] & 0x80)) {
      for (int b = 0; b < 3; ++b) A.lout[b][p] = A.lin[b][p];
      A.seg_out[p] = h;
      continue;
    }
    uint32_t lo = A.F.node_lo[h], lc = A.F.node_lc[h], lb = A.F.node_lb[h];
    for (int b = 0; b < 3; ++b) {
      bool f = (fl[r] >> b) & 1u;
      uint32_t F = run[b] - lb;  // left flags in [lo, p)
      uint32_t np = f ? lo + F : lo + lc + ((uint32_t)p - lo - F);
      A.lout[b][np] = A.lin[b][p];
      if (b == 0) A.seg_out[np] = kd_child(A.F, h, f);
      if (f) run[b] += 1u;
    }
  }
}

// seg[p] = root of p's tree (or KD_LEAF for a tree without internal nodes).
__global__ void k_seg_init(KdForest F, uint32_t* __restrict__ seg) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= F.n) return;
  if (F.point_tree) seg[p] = F.tree_root[F.point_tree[p]];
  else seg[p] = F.nnodes > 0 ? 0u : KD_LEAF;
}

// keys[i] = tree of the point vals[i] (the major key of the per-tree order).
__global__ void k_tree_keys(const uint32_t* __restrict__ point_tree, const uint32_t* __restrict__ vals, uint32_t n,
                            uint32_t* __restrict__ keys) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keys[i] = point_tree[vals[i]];
}

// Leaf order: (x, y, z, bits(tree-local photon index)); F's coordinates are
// in renumbered order, orig maps an id to its photon index.
__global__ void k_tree_points(KdForest F, const uint32_t* __restrict__ l0, const uint32_t* __restrict__ orig) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= F.n) return;
  uint32_t j = l0[p];
  uint32_t pb = F.point_tree ? F.tree_pbase[F.point_tree[p]] : 0u;
  F.tpts[p] = make_float4(F.x[j], F.y[j], F.z[j], u2f(orig[j] - pb));
}

size_t kdforest_temp_words(uint32_t n, uint32_t nnodes) {
  size_t nb = ((size_t)n + KB_TILE - 1) / KB_TILE;
  size_t w = 0;
  w += 6 * (size_t)n;      // lists (double-buffered)
  w += 2 * (size_t)n;      // seg, seg2
  w += 2 * (size_t)n;      // keys, tree keys
  w += 3 * nb + 3 * nb;    // blockcnt + blockoff
  w += scan_temp_words(3 * nb);
  w += radix_temp_words(n);
  w += 2 * (size_t)nnodes + 8;  // node_med, node_axis
  w += ((size_t)n + 3) / 4;     // per-position flag bytes of a level
  w += 5 * (size_t)n;           // renumbering: orig, inv, coordinates in the new order
  return w + 64;
}

hipError_t kd_build_forest(const KdForest& F, const uint32_t* level_off, uint32_t nlevels, uint32_t* temp,
                           hipStream_t st) {
  const uint32_t n = F.n;
  if (n == 0) return hipSuccess;
  const size_t nb = ((size_t)n + KB_TILE - 1) / KB_TILE;
  uint32_t* p = temp;
  uint32_t* lists[3];
  uint32_t* lists2[3];
  for (int a = 0; a < 3; ++a) { lists[a] = p; p += n; }
  for (int a = 0; a < 3; ++a) { lists2[a] = p; p += n; }
  uint32_t* seg = p; p += n;
  uint32_t* seg2 = p; p += n;
  uint32_t* keys = p; p += n;
  uint32_t* tkeys = p; p += n;
  uint32_t* blockcnt = p; p += 3 * nb;
  uint32_t* blockoff = p; p += 3 * nb;
  uint32_t* scantemp = p; p += scan_temp_words(3 * nb);
  uint32_t* sorttemp = p; p += radix_temp_words(n);
  uint32_t* node_med = p; p += F.nnodes + 1;
  uint32_t* node_axis = p; p += F.nnodes + 1;
  uint8_t* pflags = (uint8_t*)p; p += ((size_t)n + 3) / 4;
  uint32_t* orig = p; p += n;
  uint32_t* inv = p; p += n;
  float* xr = (float*)p; p += n;
  float* yr = (float*)p; p += n;
  float* zr = (float*)p; p += n;

  hipError_t e;
  const float* C[3] = {F.x, F.y, F.z};
  unsigned gb = (n + 255) / 256;
  // renumbering in Morton order (within each tree: a stable pass on the tree id)
  hipLaunchKernelGGL(k_morton_keys, dim3(gb), dim3(256), 0, st, F.x, F.y, F.z, n, keys, orig);
  if ((e = radix_sort_pairs(keys, orig, n, sorttemp, st)) != hipSuccess) return e;
  if (F.ntrees > 1) {
    hipLaunchKernelGGL(k_tree_keys, dim3(gb), dim3(256), 0, st, F.point_tree, orig, n, tkeys);
    if ((e = radix_pass(tkeys, orig, n, 0, sorttemp, keys, inv, st)) != hipSuccess) return e;
    uint32_t* t = orig; orig = inv; inv = t;
  }
  hipLaunchKernelGGL(k_renumber, dim3(gb), dim3(256), 0, st, (const uint32_t*)orig, n, F.x, F.y, F.z, inv, xr, yr, zr);
  for (int a = 0; a < 3; ++a) {
    hipLaunchKernelGGL(k_make_keys, dim3(gb), dim3(256), 0, st, C[a], n, keys, lists[a]);
    if ((e = radix_sort_pairs(keys, lists[a], n, sorttemp, st)) != hipSuccess) return e;
    if (F.ntrees > 1) {
      // stable pass on the tree id: per-tree lists in (key, index) order
      hipLaunchKernelGGL(k_tree_keys, dim3(gb), dim3(256), 0, st, F.point_tree, lists[a], n, tkeys);
      if ((e = radix_pass(tkeys, lists[a], n, 0, sorttemp, keys, lists2[a], st)) != hipSuccess) return e;
      uint32_t* t = lists[a]; lists[a] = lists2[a]; lists2[a] = t;
    }
    hipLaunchKernelGGL(k_map_ids, dim3(gb), dim3(256), 0, st, lists[a], n, (const uint32_t*)inv);
  }
  // from here on the build reads the coordinates in renumbered order
  KdForest R = F;
  R.x = xr; R.y = yr; R.z = zr;
  hipLaunchKernelGGL(k_seg_init, dim3(gb), dim3(256), 0, st, R, seg);
  for (uint32_t level = 0; level < nlevels; ++level) {
    const uint32_t first = level_off[level], count = level_off[level + 1] - level_off[level];
    if (count == 0) continue;
    hipLaunchKernelGGL(k_level_nodes, dim3((count + 255) / 256), dim3(256), 0, st, R, lists[0], lists[1], lists[2],
                       first, count, node_med, node_axis);
    LevelArgs A;
    A.F = R;
    A.orig = orig;
    for (int b = 0; b < 3; ++b) { A.lin[b] = lists[b]; A.lout[b] = lists2[b]; }
    A.seg = seg; A.seg_out = seg2;
    A.node_med = node_med; A.node_axis = node_axis;
    A.nblocks = (uint32_t)nb; A.blockcnt = blockcnt;
    A.flags = pflags;
    hipLaunchKernelGGL(k_level_count, dim3((unsigned)nb), dim3(KB_THREADS), 0, st, A);
    if ((e = exclusive_scan(blockcnt, blockoff, 3 * nb, scantemp, nullptr, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_level_scatter, dim3((unsigned)nb), dim3(KB_THREADS), 0, st, A, (const uint32_t*)blockoff);
    for (int b = 0; b < 3; ++b) { uint32_t* t = lists[b]; lists[b] = lists2[b]; lists2[b] = t; }
    uint32_t* t = seg; seg = seg2; seg2 = t;
  }
  hipLaunchKernelGGL(k_tree_points, dim3(gb), dim3(256), 0, st, R, (const uint32_t*)lists[0], (const uint32_t*)orig);
  return hipGetLastError();
}

size_t kdbuild_temp_words(uint32_t n, uint32_t L) {
  uint32_t nl = kd_num_leaves(n, L);
  size_t ni = nl > 0 ? nl - 1 : 0;
  return kdforest_temp_words(n, (uint32_t)ni);
}

size_t kd_shape_table_words(uint32_t n, uint32_t L) {
  uint32_t nl = kd_num_leaves(n, L);
  size_t ni = nl > 0 ? nl - 1 : 0;
  return 4 * (ni + 1);
}

// Single tree: m.tpts, m.nodes, m.nrange.  temp must hold
// kdbuild_temp_words(n, L) words; `tables` holds the device shape tables of
// the nl-1 internal nodes (heap order = level order), lo, hi, lc and lb at a
// stride of nl words.
hipError_t kd_build_impl(DevMap& m, uint32_t* temp, const uint32_t* tables, hipStream_t st) {
  const uint32_t n = m.n;
  const uint32_t nl = kd_num_leaves(n, m.L);
  m.nl = nl;
  if (n == 0) {
    m.built = true;
    return hipSuccess;
  }
  const size_t ni = nl - 1;
  const uint32_t* node_lo = tables;
  const uint32_t* node_hi = node_lo + ni + 1;
  const uint32_t* node_lc = node_hi + ni + 1;
  const uint32_t* node_lb = node_lc + ni + 1;
  hipError_t e;
  KdForest F;
  memset(&F, 0, sizeof(F));
  F.n = n;
  F.x = m.x; F.y = m.y; F.z = m.z;
  F.tpts = m.tpts;
  F.nodes = m.nodes;
  F.nrange = m.nrange;
  F.ntrees = 1;
  F.nnodes = (uint32_t)ni;
  F.node_lo = node_lo; F.node_hi = node_hi; F.node_lc = node_lc; F.node_lb = node_lb;
  std::vector<uint32_t> off(1, 0u);
  for (uint64_t f = 0, w = 1; f < ni; f = 2 * f + 1, w *= 2) off.push_back((uint32_t)std::min<uint64_t>(f + w, ni));
  if ((e = kd_build_forest(F, off.data(), (uint32_t)off.size() - 1u, temp, st)) != hipSuccess) return e;
  if ((e = kd_node_boxes(m, st)) != hipSuccess) return e;
  if ((e = kd_tree_payload(m, st)) != hipSuccess) return e;
  m.built = true;
  return hipSuccess;
}

// Leaf h = nl - 1 + i: the bounds of its points, found from its parent's
// range and left count.
__global__ void k_leaf_boxes(const float4* tpts, const KdNode* nodes, const uint2* nrange, uint32_t n, uint32_t nl,
                             float4* bmin, float4* bmax) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nl) return;
  const uint32_t h = nl - 1u + i;
  uint32_t lo = 0u, hi = n;
  if (h > 0u) {
    const uint32_t p = (h - 1u) >> 1;
    const uint2 r = nrange[p];
    const uint32_t mid = r.x + kd_left_count(nodes[p].meta);
    if (h == 2u * p + 1u) {
      lo = r.x;
      hi = mid;
    } else {
      lo = mid;
      hi = r.y;
    }
  }
  float4 mn = make_float4(INFINITY, INFINITY, INFINITY, 0.0f), mx = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
  for (uint32_t j = lo; j < hi; ++j) {
    const float4 q = tpts[j];
    mn.x = fminf(mn.x, q.x); mn.y = fminf(mn.y, q.y); mn.z = fminf(mn.z, q.z);
    mx.x = fmaxf(mx.x, q.x); mx.y = fmaxf(mx.y, q.y); mx.z = fmaxf(mx.z, q.z);
  }
  bmin[h] = mn;
  bmax[h] = mx;
}

// Internal nodes [first, first + count) of one level: the union of their
// children's boxes (the level below is done).
__global__ void k_union_boxes(uint32_t first, uint32_t count, float4* bmin, float4* bmax) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t h = first + i, l = 2u * h + 1u, r = 2u * h + 2u;
  const float4 a = bmin[l], b = bmin[r], c = bmax[l], d = bmax[r];
  bmin[h] = make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), 0.0f);
  bmax[h] = make_float4(fmaxf(c.x, d.x), fmaxf(c.y, d.y), fmaxf(c.z, d.z), 0.0f);
}

hipError_t kd_node_boxes(const DevMap& m, hipStream_t st) {
  if (!m.bmin || !m.bmax || m.n == 0 || m.nl == 0) return hipSuccess;
  const uint32_t nl = m.nl, ni = nl - 1u;
  hipLaunchKernelGGL(k_leaf_boxes, dim3((nl + 255u) / 256u), dim3(256), 0, st, (const float4*)m.tpts,
                     (const KdNode*)m.nodes, (const uint2*)m.nrange, m.n, nl, m.bmin, m.bmax);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || ni == 0) return e;
  // internal levels from the deepest up: level d holds heap nodes [2^d - 1, 2^(d+1) - 2]
  int D = 0;
  while ((2ull << D) - 1ull < (uint64_t)ni) ++D;  // deepest level with an internal node
  for (int d = D; d >= 0; --d) {
    const uint32_t first = (1u << d) - 1u;
    const uint32_t last = std::min<uint32_t>((2u << d) - 2u, ni - 1u);
    const uint32_t count = last - first + 1u;
    hipLaunchKernelGGL(k_union_boxes, dim3((count + 255u) / 256u), dim3(256), 0, st, first, count, m.bmin, m.bmax);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

// tpayload[i] = payload[record index of tpts[i]] (16 B or 32 B per record).
__global__ void k_tree_payload(const float4* tpts, const uint4* payload, uint4* tpayload, uint32_t n, int words) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t r = __float_as_uint(tpts[i].w);
  for (int w = 0; w < words; ++w) tpayload[(size_t)words * i + w] = payload[(size_t)words * r + w];
}

hipError_t kd_tree_payload(const DevMap& m, hipStream_t st) {
  if (!m.tpayload || m.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tree_payload, dim3((m.n + 255u) / 256u), dim3(256), 0, st, (const float4*)m.tpts,
                     (const uint4*)m.payload, (uint4*)m.tpayload, m.n, m.bf16 ? 1 : 2);
  return hipGetLastError();
}

// ---------------------------------------------------------------- forests
// Tree t of entry i, for a prefix table with pref[t] <= i < pref[t + 1].
__device__ __forceinline__ uint32_t forest_tree_of(const uint32_t* pref, uint32_t nt, uint32_t i) {
  uint32_t lo = 0u, hi = nt;
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pref[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void box_of_points(const float4* tpts, uint32_t lo, uint32_t hi, float4& mn, float4& mx) {
  mn = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
  mx = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
  for (uint32_t j = lo; j < hi; ++j) {
    const float4 q = tpts[j];
    mn.x = fminf(mn.x, q.x); mn.y = fminf(mn.y, q.y); mn.z = fminf(mn.z, q.z);
    mx.x = fmaxf(mx.x, q.x); mx.y = fmaxf(mx.y, q.y); mx.z = fmaxf(mx.z, q.z);
  }
}

// Leaf i of the forest: k_leaf_boxes in its own tree.
__global__ void k_forest_leaf_boxes(KdForestBoxes f) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.nleaves) return;
  const uint32_t t = forest_tree_of(f.leaf_pref, f.ntrees, i);
  const KdTreeDesc D = f.trees[t];
  const uint32_t h = D.nl - 1u + (i - f.leaf_pref[t]);
  uint32_t lo = 0u, hi = D.n;
  if (h > 0u) {
    const uint32_t p = (h - 1u) >> 1;
    const uint2 r = f.nrange[D.nbase + p];
    const uint32_t mid = r.x + kd_left_count(f.nodes[D.nbase + p].meta);
    if (h == 2u * p + 1u) {
      lo = r.x;
      hi = mid;
    } else {
      lo = mid;
      hi = r.y;
    }
  }
  float4 mn, mx;
  box_of_points(f.tpts + D.pbase, lo, hi, mn, mx);
  f.bmin[D.bbase + h] = mn;
  f.bmax[D.bbase + h] = mx;
}

// Internal nodes at level d (heap nodes 2^d - 1 ..) of every tree.
__global__ void k_forest_union_boxes(KdForestBoxes f, uint32_t d, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t* pref = f.level_pref + (size_t)d * (f.ntrees + 1u);
  const uint32_t t = forest_tree_of(pref, f.ntrees, i);
  const uint32_t bb = f.trees[t].bbase;
  const uint32_t h = (1u << d) - 1u + (i - pref[t]), l = 2u * h + 1u, r = 2u * h + 2u;
  const float4 a = f.bmin[bb + l], b = f.bmin[bb + r], c = f.bmax[bb + l], e = f.bmax[bb + r];
  f.bmin[bb + h] = make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), 0.0f);
  f.bmax[bb + h] = make_float4(fmaxf(c.x, e.x), fmaxf(c.y, e.y), fmaxf(c.z, e.z), 0.0f);
}

hipError_t kd_forest_boxes(const KdForestBoxes& f, const uint32_t* level_total, hipStream_t st) {
  if (f.nleaves == 0) return hipSuccess;
  hipLaunchKernelGGL(k_forest_leaf_boxes, dim3((f.nleaves + 255u) / 256u), dim3(256), 0, st, f);
  hipError_t e = hipGetLastError();
  for (uint32_t d = f.depth; d-- > 0u && e == hipSuccess;) {
    const uint32_t count = level_total[d];
    if (count == 0) continue;
    hipLaunchKernelGGL(k_forest_union_boxes, dim3((count + 255u) / 256u), dim3(256), 0, st, f, d, count);
    e = hipGetLastError();
  }
  return e;
}

__global__ void k_forest_payload(const float4* tpts, const uint32_t* point_tree, const uint32_t* tree_pbase,
                                 const uint4* payload, uint4* tpayload, uint32_t n, int words) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t r = tree_pbase[point_tree[i]] + __float_as_uint(tpts[i].w);
  for (int w = 0; w < words; ++w) tpayload[(size_t)words * i + w] = payload[(size_t)words * r + w];
}

hipError_t kd_forest_payload(const float4* tpts, const uint32_t* point_tree, const uint32_t* tree_pbase,
                             const void* payload, void* tpayload, uint32_t n, int bf16, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_forest_payload, dim3((n + 255u) / 256u), dim3(256), 0, st, tpts, point_tree, tree_pbase,
                     (const uint4*)payload, (uint4*)tpayload, n, bf16 ? 1 : 2);
  return hipGetLastError();
}

}  // namespace pm
