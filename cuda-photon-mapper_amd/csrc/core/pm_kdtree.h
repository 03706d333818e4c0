// pm_kdtree.h -- shape of the canonical left-balanced kd-tree over photon
// positions, shared by the GPU builder, the GPU gather and the host oracle.
//
// The reference has no spatial structure beyond getVoxelCoordinates
// (photonMappingKernel.cu:260-267, a fixed 32^3 grid); this tree replaces it.
//
// Shape: n points, leaf size L, nl = ceil(n/L) leaves.  The tree is the
// complete binary tree with nl leaves in heap order: internal nodes are
// 0 .. nl-2 (children 2h+1, 2h+2), leaves are nl-1 .. 2nl-2.  Points are
// assigned to leaves in in-order: every leaf holds L points except the
// rightmost one, which holds n - (nl-1) L.  A node covering points [lo, hi)
// therefore sends the first L * leaves(2h+1) of them left.
//
// Split rule at an internal node: axis = largest bounding-box extent of its
// points (ties -> lowest axis); points ordered by (coordinate key, photon
// index); the first point of the right part is the median and its coordinate
// is the split value.  Inside a leaf points are ordered by (x key, index).
#pragma once

#include "pm_core.h"

namespace pm {

PM_HD uint32_t kd_num_leaves(uint32_t n, uint32_t L) { return n == 0 ? 0u : (n + L - 1u) / L; }

// Number of leaves in the subtree of heap node c for a tree with nl leaves.
PM_HD uint32_t kd_leaves_under(uint32_t c, uint32_t nl) {
  uint64_t first = c, last = c, cnt = 0;
  const uint64_t lo_leaf = (uint64_t)nl - 1u, hi_leaf = 2ull * nl - 2u;
  while (first <= hi_leaf) {
    uint64_t a = first > lo_leaf ? first : lo_leaf;
    uint64_t b = last < hi_leaf ? last : hi_leaf;
    if (b >= a) cnt += b - a + 1u;
    first = 2u * first + 1u;
    last = 2u * last + 2u;
  }
  return (uint32_t)cnt;
}

// Internal node record: split value and (axis | left point count << 2).
struct KdNode {
  float split;
  uint32_t meta;
};
PM_HD int kd_axis(uint32_t meta) { return (int)(meta & 3u); }
PM_HD uint32_t kd_left_count(uint32_t meta) { return meta >> 2; }

// Point range [lo, hi) of heap node h (walks down from the root).
PM_HD void kd_node_range(uint32_t h, uint32_t n, uint32_t L, uint32_t* lo_out, uint32_t* hi_out) {
  uint32_t nl = kd_num_leaves(n, L);
  // Path from root to h.
  uint32_t path[40];
  int depth = 0;
  uint32_t c = h;
  while (c > 0) {
    path[depth++] = c;
    c = (c - 1u) / 2u;
  }
  uint32_t lo = 0, hi = n;
  for (int i = depth - 1; i >= 0; --i) {
    uint32_t child = path[i];
    uint32_t parent = (child - 1u) / 2u;
    uint32_t lc = L * kd_leaves_under(2u * parent + 1u, nl);
    if (child == 2u * parent + 1u) hi = lo + lc;
    else lo = lo + lc;
  }
  *lo_out = lo;
  *hi_out = hi;
}

// kNN ordering key: (FP32 bits of the squared distance, photon index).
PM_HD uint64_t knn_key(float d2, uint32_t idx) { return ((uint64_t)f2u(d2) << 32) | (uint64_t)idx; }

// Squared distance in fixed order, no fused multiply-add.
PM_HD float dist2(float qx, float qy, float qz, float px, float py, float pz) {
  float dx = px - qx, dy = py - qy, dz = pz - qz;
  return (dx * dx + dy * dy) + dz * dz;
}

}  // namespace pm
