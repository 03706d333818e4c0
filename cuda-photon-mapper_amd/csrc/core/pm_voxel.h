// pm_voxel.h -- the voxel-grid photon estimate: the reference's method, kept as
// a compatibility mode beside the k-NN gather.  Shared per-thread code (GPU
// kernels and host oracle).
//
// Reference counterparts (photonMappingKernel.cu):
//   the 32^3 grid `photons` and its voxel coordinates        :76, :260-267
//   storePhoton / storeVolumePhoton (photon into its voxel)  :1147-1183
//   integrate / gatherPhotons: the window [v-3, v+3) of the
//     wall's boundary layer (wall 0: i = 31, 2: i = 0, 1: j = 0,
//     3: j = 31, 4: k = 31), clamped to the grid             :314-389, :586-600
//   bilinearInterpolate / interpolateEnergy (interpolateFlag):
//     window sums at the 4 nearest voxel centres, blended    :442-583
//   integrateVolumePhotons: the window [v-3, v+3)^3 clamped
//     to the interior [1, 31)                                :831-870
// Kept from the reference: the 32^3 grid over the room, the window shapes and
// their clamping, the boundary layers, the bilinear blend.  Changed: the grid
// holds photon flux in 64-bit fixed point (2^-24 units of a record's relative
// power), added with integer atomics, so its sums are exact and the same in
// any order (the reference's float += from concurrent threads races); surface
// and volume records go to separate grids; the estimates are physical
// (window flux over the window's wall area, times albedo / pi; window flux
// over its volume, times the isotropic phase 1 / 4 pi) instead of fixed
// weights; window sums come from summed-volume tables, 8 lookups each.
// Hits on spheres get no surface estimate, as in the reference (:343).
#pragma once

#include "pm_scene.h"

namespace pm {

constexpr int kVox = 32;              // voxels per axis over the unit room
constexpr int kVoxHalf = 3;           // window [v - 3, v + 3)
constexpr int kSat = kVox + 1;        // summed-volume table side
constexpr double kVoxFix = 16777216.0;  // fixed-point units per unit of record power (2^24)

// Voxel of coordinate p in [0, 1] (truncation, clamped to the grid).
PM_HD int vox_of(float p) {
  int v = (int)(p * (float)kVox);
  return v < 0 ? 0 : (v > kVox - 1 ? kVox - 1 : v);
}
template <class T> PM_HD int vox_of_t(T p) { return vox_of((float)p); }

// A record's power in fixed point (power >= 0): round to nearest.
PM_HD uint64_t vox_fix(float p) { return (uint64_t)((double)p * kVoxFix + 0.5); }

// Grid cell (x, y, z), channel ch, of a [kVox^3][3] grid.
PM_HD size_t vox_cell(int x, int y, int z) { return ((size_t)z * kVox + (size_t)y) * kVox + (size_t)x; }
// Summed-volume table entry: S(x, y, z) = sum of cells [0, x) x [0, y) x [0, z).
PM_HD size_t sat_idx(int x, int y, int z) { return ((size_t)z * kSat + (size_t)y) * kSat + (size_t)x; }

// Sum over the box [x0, x1) x [y0, y1) x [z0, z1) of channel ch (exact; the
// modular arithmetic of the inclusion-exclusion cancels).
PM_HD uint64_t sat_box(const uint64_t* S, int ch, int x0, int x1, int y0, int y1, int z0, int z1) {
  auto at = [&](int x, int y, int z) { return S[3 * sat_idx(x, y, z) + (size_t)ch]; };
  return ((at(x1, y1, z1) - at(x0, y1, z1)) - (at(x1, y0, z1) - at(x0, y0, z1))) -
         ((at(x1, y1, z0) - at(x0, y1, z0)) - (at(x1, y0, z0) - at(x0, y0, z0)));
}

PM_HD int vclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

struct VoxelGrids {
  const uint64_t* sat_s;  // surface records, summed-volume table [kSat^3][3]
  const uint64_t* sat_v;  // volume records
  float scale_s[3];       // light power / photons emitted, per channel
  float scale_v[3];
  int interp;             // the reference's interpolateFlag: bilinear blend of 4 windows
};

// Surface flux density (W / m^2) of the window centred on cell (cb, cc) of the
// boundary layer `layer` of axis a (in-plane axes b, c), per channel.
PM_HD void vox_surface_window(const VoxelGrids& G, int a, int layer, int cb, int cc, double out[3]) {
  const int b0 = vclamp(cb - kVoxHalf, 0, kVox), b1 = vclamp(cb + kVoxHalf, 0, kVox);
  const int c0 = vclamp(cc - kVoxHalf, 0, kVox), c1 = vclamp(cc + kVoxHalf, 0, kVox);
  const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
  int lo[3], hi[3];
  lo[a] = layer; hi[a] = layer + 1;
  lo[b] = b0; hi[b] = b1;
  lo[c] = c0; hi[c] = c1;
  const double area = (double)((b1 - b0) * (c1 - c0)) / (double)(kVox * kVox);
  for (int ch = 0; ch < 3; ++ch) {
    const uint64_t q = sat_box(G.sat_s, ch, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]);
    out[ch] = (double)q / kVoxFix * (double)G.scale_s[ch] / area;
  }
}

template <class T>
struct VoxelGather {
  VoxelGrids G;
  uint32_t nsurf, nvol;
  // the last volume window: consecutive march points mostly share a voxel,
  // and the estimate is a function of the voxel alone
  int last_cell = -1;
  V3<T> last_vol;

  // Outgoing radiance of a diffuse wall point x with normal n (toward the
  // viewer): albedo / pi times the window's flux density (bilinear blend of
  // the 4 windows around x when interpolating).  0 on spheres.
  PM_HD V3<T> surface(V3<T> x, V3<T> n, V3<T> albedo, V3<T> = V3<T>(), V3<T> = V3<T>()) {
    ++nsurf;
    const T ax = n.x < T(0) ? -n.x : n.x, ay = n.y < T(0) ? -n.y : n.y, az = n.z < T(0) ? -n.z : n.z;
    // a wall's normal is one axis (the other components exactly 0)
    const T one = T(1), zero = T(0);
    const int a = (ax == one && ay == zero && az == zero)   ? 0
                  : (ay == one && ax == zero && az == zero) ? 1
                  : (az == one && ax == zero && ay == zero) ? 2
                                                            : -1;
    if (a < 0) return v3<T>(T(0), T(0), T(0));
    const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
    const T pa = comp(x, a), pb = comp(x, b), pc = comp(x, c);
    const int layer = pa > T(0.5) ? kVox - 1 : 0;
    double e[3];
    if (!G.interp) {
      vox_surface_window(G, a, layer, vox_of_t(pb), vox_of_t(pc), e);
    } else {
      // voxel centres at (i + 0.5) / 32: blend the windows of the 4 nearest
      const double ub = (double)pb * kVox - 0.5, uc = (double)pc * kVox - 0.5;
      const double fb0 = ub < 0.0 ? -1.0 : (double)(long long)ub, fc0 = uc < 0.0 ? -1.0 : (double)(long long)uc;
      double fb = ub - fb0, fc = uc - fc0;
      fb = fb < 0.0 ? 0.0 : (fb > 1.0 ? 1.0 : fb);
      fc = fc < 0.0 ? 0.0 : (fc > 1.0 ? 1.0 : fc);
      const int ib = (int)fb0, ic = (int)fc0;
      double w00[3], w10[3], w01[3], w11[3];
      vox_surface_window(G, a, layer, vclamp(ib, 0, kVox - 1), vclamp(ic, 0, kVox - 1), w00);
      vox_surface_window(G, a, layer, vclamp(ib + 1, 0, kVox - 1), vclamp(ic, 0, kVox - 1), w10);
      vox_surface_window(G, a, layer, vclamp(ib, 0, kVox - 1), vclamp(ic + 1, 0, kVox - 1), w01);
      vox_surface_window(G, a, layer, vclamp(ib + 1, 0, kVox - 1), vclamp(ic + 1, 0, kVox - 1), w11);
      for (int ch = 0; ch < 3; ++ch) {
        const double lo = w00[ch] + fb * (w10[ch] - w00[ch]);
        const double hi = w01[ch] + fb * (w11[ch] - w01[ch]);
        e[ch] = lo + fc * (hi - lo);
      }
    }
    const double inv_pi = 0.318309886183790671537767526745;
    return v3<T>(T((double)albedo.x * inv_pi * e[0]), T((double)albedo.y * inv_pi * e[1]),
                 T((double)albedo.z * inv_pi * e[2]));
  }

  // In-scattered radiance source at x: the interior window's flux density
  // per unit volume times the isotropic phase 1 / (4 pi).
  PM_HD V3<T> volume(V3<T> x, V3<T>, V3<T> = V3<T>()) {
    ++nvol;
    const int vx = vox_of_t(x.x), vy = vox_of_t(x.y), vz = vox_of_t(x.z);
    const int cell = (vz * kVox + vy) * kVox + vx;
    if (cell == last_cell) return last_vol;
    last_cell = cell;
    const int x0 = vclamp(vx - kVoxHalf, 1, kVox - 1), x1 = vclamp(vx + kVoxHalf, 1, kVox - 1);
    const int y0 = vclamp(vy - kVoxHalf, 1, kVox - 1), y1 = vclamp(vy + kVoxHalf, 1, kVox - 1);
    const int z0 = vclamp(vz - kVoxHalf, 1, kVox - 1), z1 = vclamp(vz + kVoxHalf, 1, kVox - 1);
    const int cells = (x1 - x0) * (y1 - y0) * (z1 - z0);
    if (cells <= 0) return last_vol = v3<T>(T(0), T(0), T(0));
    const double vol = (double)cells / (double)(kVox * kVox * kVox);
    const double inv4pi = 0.0795774715459476678844418816863;
    double s[3];
    for (int ch = 0; ch < 3; ++ch) {
      const uint64_t q = sat_box(G.sat_v, ch, x0, x1, y0, y1, z0, z1);
      s[ch] = (double)q / kVoxFix * (double)G.scale_v[ch] / vol * inv4pi;
    }
    return last_vol = v3<T>(T(s[0]), T(s[1]), T(s[2]));
  }
};

}  // namespace pm
