"""Scenes with closed-form answers, shared by the CPU pin tests
(test_oracle_pins.py) and their GPU counterparts (test_gpu_pins.py).

A map of volume photons uniform in a box larger than the room (density rho,
unit power, given directions) with an empty-power surface map: every radiance
query is a volume k-NN estimate, whose expectation for a Poisson point set is
rho k / (k - 1) * Phi p (the k-th neighbour's ball V_k has E[1 / V_k] =
rho / (k - 1)), and a pixel's radiance is that estimate times the march
weights sum_j ds exp(-sigma_t (t_j - t0)) of its camera ray.
"""
import math

import numpy as np

import oracle as O


def hg(c, g):
    """Henyey-Greenstein phase function of cos(theta), written from its definition."""
    return (1 - g * g) / (4 * math.pi * (1 + g * g - 2 * g * c) ** 1.5)


def sample_ts(t0, t1, xi, ds):
    """Sample positions of the march, t_j = t0 + (j + xi) ds < t1, evaluated in
    FP32 as the per-thread code does."""
    t0, t1, xi, ds = np.float32(t0), np.float32(t1), np.float32(xi), np.float32(ds)
    ts = []
    j = 0
    while True:
        t = np.float32(t0 + np.float32(np.float32(j) + xi) * ds)
        if not t < t1:
            return np.array(ts, np.float64)
        ts.append(float(t))
        j += 1


def uniform_volume_scene(k=64, n=1_000_000, g=0.0, dirs=None, seed=11):
    """Preset-1 room without spheres (homogeneous medium), 64x64 at 1 spp, and n
    volume photons uniform in [-0.1, 1.1]^3 (no query ball is truncated by the
    walls).  Returns (config, positions, directions, density)."""
    c = O.preset(1)
    c.n_spheres = 0
    c.k_volume, c.rcap_volume, c.hg_g = k, 0.25, g
    c.width, c.height = 64, 64
    rng = np.random.default_rng(seed)
    pos = (-0.1 + 1.2 * rng.random((n, 3))).astype(np.float32)
    if dirs is None:
        d = rng.normal(size=(n, 3))
        dirs = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    else:
        dirs = np.broadcast_to(np.asarray(dirs, np.float32), (n, 3)).copy()
    rho = n / 1.2 ** 3
    return c, pos, dirs, rho


def zero_surface_map(n=1000, seed=12):
    """n surface photons on the back wall with zero power: every surface
    estimate is exactly 0."""
    rng = np.random.default_rng(seed)
    pos = np.ones((n, 3), np.float32)
    pos[:, :2] = rng.random((n, 2), dtype=np.float32)
    power = np.zeros((n, 3), np.float32)
    dirs = np.zeros((n, 3), np.float32)
    dirs[:, 2] = 1.0
    return pos, power, dirs


def load_uniform(m, c, pos, dirs, n_emit):
    """Loads the scene into an oracle or a libpm context (both have
    set_records / build) and builds the trees."""
    m.set_records(1, pos, np.ones_like(pos), dirs, n_emit)
    m.set_records(0, *zero_surface_map(), n_emit)
    m.build()


def march_weights(o, c, px, py):
    """sum_j ds T_j of the pixel's march (homogeneous medium) from the open front
    face to the wall the camera ray hits, and the ray direction; o is an oracle
    of the same configuration (camera ray and intersection)."""
    st, ds = c.sigma_s + c.sigma_a, c.march_step
    ro, rd, xi = o.camera_ray(px, py, 0)
    t0 = -float(ro[2]) / float(rd[2])
    t1 = float(o.intersect(ro, rd)[0][0])
    ts = sample_ts(t0, t1, xi, ds)
    return float(np.float32(ds)) * np.exp(-st * (ts - t0)).sum(), rd


def uniform_ratio(img, o, c, rho, n_emit):
    """Per-pixel ratio of the rendered radiance (red channel) to the Poisson
    closed form rho k / (k - 1) * (Phi / n_emit) / (4 pi) * march weights (g = 0)."""
    k = c.k_volume
    S = (c.light_power[0] / n_emit) / (4 * math.pi) * rho * k / (k - 1)
    r = np.empty((c.height, c.width))
    for py in range(c.height):
        for px in range(c.width):
            W, _ = march_weights(o, c, px, py)
            r[py, px] = img[py, px, 0] / (S * W)
    return r
