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
#include "pm_kernels.h"
#include "pm_wave.h"

namespace pm {

static constexpr int KB_THREADS = 256;
static constexpr int KB_ITEMS = 4;
static constexpr int KB_TILE = KB_THREADS * KB_ITEMS;

__device__ __forceinline__ float coord_of(const float* x, const float* y, const float* z, uint32_t i, int a) {
  return a == 0 ? x[i] : (a == 1 ? y[i] : z[i]);
}

__global__ void k_make_keys(const float* __restrict__ c, uint32_t n, uint32_t* __restrict__ keys,
                            uint32_t* __restrict__ vals) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = float_key(c[i]);
  vals[i] = i;
}

// One thread per internal node of the level [f, e).
__global__ void k_level_nodes(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                              const uint32_t* __restrict__ l0, const uint32_t* __restrict__ l1,
                              const uint32_t* __restrict__ l2, uint32_t f, uint32_t e,
                              const uint32_t* __restrict__ node_lo, const uint32_t* __restrict__ node_hi,
                              const uint32_t* __restrict__ node_lc, KdNode* __restrict__ nodes,
                              uint2* __restrict__ nrange, uint32_t* __restrict__ node_med,
                              uint32_t* __restrict__ node_axis) {
  uint32_t h = f + blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= e) return;
  uint32_t lo = node_lo[h], hi = node_hi[h], lc = node_lc[h];
  const uint32_t* L[3] = {l0, l1, l2};
  float ext[3];
  for (int a = 0; a < 3; ++a) {
    float mn = coord_of(x, y, z, L[a][lo], a);
    float mx = coord_of(x, y, z, L[a][hi - 1], a);
    ext[a] = mx - mn;
  }
  int axis = 0;
  float best = ext[0];
  if (ext[1] > best) { best = ext[1]; axis = 1; }
  if (ext[2] > best) { best = ext[2]; axis = 2; }
  uint32_t med = L[axis][lo + lc];
  KdNode nd;
  nd.split = coord_of(x, y, z, med, axis);
  nd.meta = (uint32_t)axis | (lc << 2);
  nodes[h] = nd;
  nrange[h] = make_uint2(lo, hi);
  node_med[h] = med;
  node_axis[h] = (uint32_t)axis;
}

struct LevelArgs {
  const float* x;
  const float* y;
  const float* z;
  const uint32_t* lin[3];
  uint32_t* lout[3];
  const uint32_t* seg;
  uint32_t* seg_out;
  uint32_t f, e;  // internal nodes of this level
  const uint32_t* node_lo;
  const uint32_t* node_lc;
  const uint32_t* node_lb;
  const uint32_t* node_med;
  const uint32_t* node_axis;
  uint32_t n;
  uint32_t nblocks;
  uint32_t* blockcnt;  // [3][nblocks]
};

// Flag of list b at position p: does the element order before the node's
// median under (key on the node's axis, index)?  On the node's own axis the
// list is already split at the median, so the flag is "p < median position".
__device__ __forceinline__ bool part_flag(const LevelArgs& A, int b, uint32_t p, uint32_t h) {
  uint32_t a = A.node_axis[h];
  if ((uint32_t)b == a) return p < A.node_lo[h] + A.node_lc[h];
  uint32_t i = A.lin[b][p];
  uint32_t m = A.node_med[h];
  uint32_t ki = float_key(coord_of(A.x, A.y, A.z, i, (int)a));
  uint32_t km = float_key(coord_of(A.x, A.y, A.z, m, (int)a));
  return ki < km || (ki == km && i < m);
}

__global__ void __launch_bounds__(KB_THREADS) k_level_count(LevelArgs A) {
  __shared__ uint32_t lw[KB_THREADS / 64];
  uint32_t cnt[3] = {0u, 0u, 0u};
  size_t base = (size_t)blockIdx.x * KB_TILE + (size_t)threadIdx.x * KB_ITEMS;
  for (int r = 0; r < KB_ITEMS; ++r) {
    size_t p = base + r;
    if (p >= A.n) break;
    uint32_t h = A.seg[p];
    if (h < A.f || h >= A.e) continue;
    for (int b = 0; b < 3; ++b) cnt[b] += part_flag(A, b, (uint32_t)p, h) ? 1u : 0u;
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
__global__ void __launch_bounds__(KB_THREADS) k_level_scatter(LevelArgs A, const uint32_t* __restrict__ blockoff) {
  __shared__ uint32_t lw[KB_THREADS / 64];
  size_t base = (size_t)blockIdx.x * KB_TILE + (size_t)threadIdx.x * KB_ITEMS;
  uint32_t hs[KB_ITEMS];
  uint8_t fl[KB_ITEMS];  // bit b = flag of list b; bit 7 = position is in a node of this level
  uint32_t cnt[3] = {0u, 0u, 0u};
  for (int r = 0; r < KB_ITEMS; ++r) {
    size_t p = base + r;
    fl[r] = 0;
    hs[r] = 0;
    if (p >= A.n) continue;
    uint32_t h = A.seg[p];
    hs[r] = h;
    if (h < A.f || h >= A.e) continue;
    uint8_t m = 0x80;
    for (int b = 0; b < 3; ++b)
      if (part_flag(A, b, (uint32_t)p, h)) {
        m |= (uint8_t)(1u << b);
        cnt[b] += 1u;
      }
    fl[r] = m;
  }
  uint32_t run[3];
  for (int b = 0; b < 3; ++b) {
    uint32_t boff = blockoff[(size_t)b * A.nblocks + blockIdx.x] - blockoff[(size_t)b * A.nblocks];
    run[b] = boff + kb_block_excl(cnt[b], lw);
  }
  for (int r = 0; r < KB_ITEMS; ++r) {
    size_t p = base + r;
    if (p >= A.n) break;
    uint32_t h = hs[r];
    if (!(fl[r] & 0x80)) {
      for (int b = 0; b < 3; ++b) A.lout[b][p] = A.lin[b][p];
      A.seg_out[p] = h;
      continue;
    }
    uint32_t lo = A.node_lo[h], lc = A.node_lc[h], lb = A.node_lb[h];
    for (int b = 0; b < 3; ++b) {
      bool f = (fl[r] >> b) & 1u;
      uint32_t F = run[b] - lb;  // left flags in [lo, p)
      uint32_t np = f ? lo + F : lo + lc + ((uint32_t)p - lo - F);
      A.lout[b][np] = A.lin[b][p];
      if (b == 0) A.seg_out[np] = f ? 2u * h + 1u : 2u * h + 2u;
      if (f) run[b] += 1u;
    }
  }
}

__global__ void k_tree_points(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                              const uint32_t* __restrict__ l0, uint32_t n, float4* __restrict__ out) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t i = l0[p];
  out[p] = make_float4(x[i], y[i], z[i], u2f(i));
}

size_t kdbuild_temp_words(uint32_t n, uint32_t L) {
  uint32_t nl = kd_num_leaves(n, L);
  size_t ni = nl > 0 ? nl - 1 : 0;
  size_t nb = ((size_t)n + KB_TILE - 1) / KB_TILE;
  size_t w = 0;
  w += 6 * (size_t)n;      // lists (double-buffered)
  w += 2 * (size_t)n;      // seg, seg2
  w += (size_t)n;          // keys
  w += 3 * nb + 3 * nb;    // blockcnt + blockoff
  w += scan_temp_words(3 * nb);
  w += radix_temp_words(n);
  w += 6 * ni + 8;         // node_lo, hi, lc, lb, med, axis
  return w + 64;
}

// Builds m.tpts and m.nodes.  temp must hold kdbuild_temp_words(n, L) words;
// host_lo/host_hi/host_lc/host_lb are the shape tables of the nl-1 internal nodes.
hipError_t kd_build_impl(DevMap& m, uint32_t* temp, const uint32_t* host_lo, const uint32_t* host_hi,
                         const uint32_t* host_lc, const uint32_t* host_lb, hipStream_t st) {
  const uint32_t n = m.n;
  const uint32_t nl = kd_num_leaves(n, m.L);
  m.nl = nl;
  if (n == 0) {
    m.built = true;
    return hipSuccess;
  }
  const size_t ni = nl - 1;
  const size_t nb = ((size_t)n + KB_TILE - 1) / KB_TILE;
  uint32_t* p = temp;
  uint32_t* lists[3];
  uint32_t* lists2[3];
  for (int a = 0; a < 3; ++a) { lists[a] = p; p += n; }
  for (int a = 0; a < 3; ++a) { lists2[a] = p; p += n; }
  uint32_t* seg = p; p += n;
  uint32_t* seg2 = p; p += n;
  uint32_t* keys = p; p += n;
  uint32_t* blockcnt = p; p += 3 * nb;
  uint32_t* blockoff = p; p += 3 * nb;
  uint32_t* scantemp = p; p += scan_temp_words(3 * nb);
  uint32_t* sorttemp = p; p += radix_temp_words(n);
  uint32_t* node_lo = p; p += ni + 1;
  uint32_t* node_hi = p; p += ni + 1;
  uint32_t* node_lc = p; p += ni + 1;
  uint32_t* node_lb = p; p += ni + 1;
  uint32_t* node_med = p; p += ni + 1;
  uint32_t* node_axis = p; p += ni + 1;

  hipError_t e;
  const float* C[3] = {m.x, m.y, m.z};
  unsigned gb = (n + 255) / 256;
  for (int a = 0; a < 3; ++a) {
    hipLaunchKernelGGL(k_make_keys, dim3(gb), dim3(256), 0, st, C[a], n, keys, lists[a]);
    if ((e = radix_sort_pairs(keys, lists[a], n, sorttemp, st)) != hipSuccess) return e;
  }
  if (ni > 0) {
    if ((e = hipMemcpyAsync(node_lo, host_lo, 4 * ni, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(node_hi, host_hi, 4 * ni, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(node_lc, host_lc, 4 * ni, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(node_lb, host_lb, 4 * ni, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
  }
  if ((e = hipMemsetAsync(seg, 0, 4 * (size_t)n, st)) != hipSuccess) return e;
  for (uint32_t level = 0;; ++level) {
    uint64_t f = (1ull << level) - 1ull;
    if (f >= ni) break;
    uint64_t eend = (2ull << level) - 1ull;
    if (eend > ni) eend = ni;
    unsigned nbn = (unsigned)((eend - f + 255) / 256);
    hipLaunchKernelGGL(k_level_nodes, dim3(nbn), dim3(256), 0, st, m.x, m.y, m.z, lists[0], lists[1], lists[2],
                       (uint32_t)f, (uint32_t)eend, node_lo, node_hi, node_lc, m.nodes, m.nrange, node_med,
                       node_axis);
    LevelArgs A;
    A.x = m.x; A.y = m.y; A.z = m.z;
    for (int b = 0; b < 3; ++b) { A.lin[b] = lists[b]; A.lout[b] = lists2[b]; }
    A.seg = seg; A.seg_out = seg2;
    A.f = (uint32_t)f; A.e = (uint32_t)eend;
    A.node_lo = node_lo; A.node_lc = node_lc; A.node_lb = node_lb; A.node_med = node_med; A.node_axis = node_axis;
    A.n = n; A.nblocks = (uint32_t)nb; A.blockcnt = blockcnt;
    hipLaunchKernelGGL(k_level_count, dim3((unsigned)nb), dim3(KB_THREADS), 0, st, A);
    if ((e = exclusive_scan(blockcnt, blockoff, 3 * nb, scantemp, nullptr, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_level_scatter, dim3((unsigned)nb), dim3(KB_THREADS), 0, st, A, (const uint32_t*)blockoff);
    for (int b = 0; b < 3; ++b) { uint32_t* t = lists[b]; lists[b] = lists2[b]; lists2[b] = t; }
    uint32_t* t = seg; seg = seg2; seg2 = t;
  }
  hipLaunchKernelGGL(k_tree_points, dim3(gb), dim3(256), 0, st, m.x, m.y, m.z, lists[0], n, m.tpts);
  m.built = true;
  return hipGetLastError();
}

}  // namespace pm
