"""Pins of the per-thread core shared by the oracle and the gfx950 kernels.

Each check is against something independent of the code under test: published
known-answer vectors and rocRAND's engine (Philox), libm in FP64 (own sin, cos,
log, exp), closed forms (intersections, Fresnel, transmittance) and moments of
the sampled distributions.  Runs on the CPU.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
import pmamd

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------- Philox
def test_philox_random123_kat():
    # Known-answer vectors of Philox4x32-10 (Salmon et al., SC'11; Random123 kat_vectors).
    ctr = [[0, 0, 0, 0], [0xFFFFFFFF] * 4, [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344]]
    key = [[0, 0], [0xFFFFFFFF] * 2, [0xA4093822, 0x299F31D0]]
    want = [[0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8],
            [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD],
            [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]]
    got = O.philox(np.array(ctr, np.uint32), np.array(key, np.uint32))
    assert got.tolist() == want


def test_philox_matches_rocrand_engine():
    # Fixture written by tests/golden/gen_philox_rocrand.sh from rocRAND's host engine.
    cases = json.load(open(os.path.join(HERE, "golden", "philox_rocrand.json")))
    ctr = np.array([c["counter"] for c in cases], np.uint32)
    key = np.array([c["key"] for c in cases], np.uint32)
    got = O.philox(ctr, key)
    assert got.tolist() == [c["out"] for c in cases]


def test_philox_bits_uniform():
    ctr = np.zeros((200000, 4), np.uint32)
    ctr[:, 0] = np.arange(200000)
    ctr[:, 3] = 1
    out = O.philox(ctr, np.tile(np.array([1, 0], np.uint32), (200000, 1)))
    bits = np.unpackbits(out.view(np.uint8))
    assert abs(bits.mean() - 0.5) < 1e-3
    u = (out[:, 0] >> 8).astype(np.float64) / 2 ** 24
    hist = np.histogram(u, bins=64, range=(0, 1))[0]
    chi2 = ((hist - hist.mean()) ** 2 / hist.mean()).sum()
    assert chi2 < 120  # 63 dof


# ---------------------------------------------------------------- own transcendentals
def ulp_err(got, ref):
    ref32 = ref.astype(np.float32)
    spacing = np.spacing(np.abs(ref32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref) / spacing


def test_sincos_turns_accuracy():
    u = np.concatenate([np.arange(0, 2 ** 24, 97, dtype=np.float64) / 2 ** 24,
                        np.array([0.0, 0.125, 0.25, 0.375, 0.5, 0.75, 1 - 2 ** -24])]).astype(np.float32)
    s, c = O.math(0, u), O.math(1, u)
    ref_s = np.sin(2 * np.pi * u.astype(np.float64))
    ref_c = np.cos(2 * np.pi * u.astype(np.float64))
    # absolute error bound (values near zero crossings): 4 ulp of 1.0
    assert np.abs(s - ref_s).max() < 4 * 2 ** -24
    assert np.abs(c - ref_c).max() < 4 * 2 ** -24
    assert s[u == 0.25][0] == 1.0 and c[u == 0.0][0] == 1.0 and s[u == 0.0][0] == 0.0


def test_log_accuracy():
    x = np.concatenate([1.0 - np.arange(0, 2 ** 24, 13, dtype=np.float64) / 2 ** 24,
                        np.geomspace(1e-38, 1e38, 20000), np.array([1.0, 2.0, 0.5, 2 ** -24])]).astype(np.float32)
    y = O.math(2, x)
    ref = np.log(x.astype(np.float64))
    err = np.abs(y - ref)
    tol = 3 * np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64) + 1e-45
    assert (err <= np.maximum(tol, 2 * 2 ** -24)).all()
    assert y[x == 1.0][0] == 0.0


def test_exp_accuracy():
    x = np.concatenate([np.linspace(-100, 88, 400001), np.array([0.0, 1.0, -1.0])]).astype(np.float32)
    y = O.math(3, x)
    ref = np.exp(x.astype(np.float64))
    normal = ref > 1.2e-38
    assert ulp_err(y[normal], ref[normal]).max() <= 3.0
    assert y[x == 0.0][0] == 1.0
    assert np.abs(y[~normal] - ref[~normal]).max() < 2e-45 * 4


# ---------------------------------------------------------------- sampling
def uniforms(n, seed=3):
    r = np.random.default_rng(seed)
    return (np.floor(r.random((n, 2)) * 2 ** 24) / 2 ** 24).astype(np.float32)


def test_sample_sphere_moments():
    d = O.sample(0, [], uniforms(400000)).astype(np.float64)
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-6
    assert np.abs(d.mean(0)).max() < 5e-3
    np.testing.assert_allclose((d ** 2).mean(0), [1 / 3] * 3, atol=3e-3)


@pytest.mark.parametrize("n", [[0, 0, 1], [0, 0, -1], [1, 0, 0], [0.6, -0.8, 0.0], [0.48, 0.6, -0.64]])
def test_sample_cosine_moments(n):
    n = np.array(n, np.float64)
    d = O.sample(1, n.astype(np.float32), uniforms(400000)).astype(np.float64)
    c = d @ n
    assert c.min() >= -1e-6
    assert abs(c.mean() - 2 / 3) < 3e-3          # E[cos] = 2/3 for a cosine-weighted lobe
    assert abs((c ** 2).mean() - 1 / 2) < 3e-3   # E[cos^2] = 1/2
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-5


@pytest.mark.parametrize("g", [0.0, 0.6, -0.3, 0.9])
def test_sample_hg_mean_cosine(g):
    d0 = np.array([0.0, 0.6, 0.8])
    d = O.sample(2, [*d0.astype(np.float32), g], uniforms(400000)).astype(np.float64)
    c = d @ d0
    assert abs(c.mean() - g) < 4e-3  # E[cos theta] = g for Henyey-Greenstein


@pytest.mark.parametrize("g", [0.0, 0.6, -0.5])
def test_phase_hg_normalised(g):
    mu = np.linspace(-1, 1, 20001)
    p = np.array([O.phase_hg(float(m), g) for m in mu])
    integral = 2 * np.pi * np.trapezoid(p, mu)
    assert abs(integral - 1) < 1e-4


@pytest.mark.parametrize("g", [-0.95, -0.6, -0.1, 0.0, 0.3, 0.6, 0.9])
def test_phase_hg_values(g):
    # The FP32 phase takes den = 1 + g^2 - 2 g mu and 1 - g^2 in FP32, then
    # den^-3/2 by Newton steps.  Against the closed form in FP64 on those
    # FP32 values: within 6e-7 plus the two final products' roundings.
    f = np.float32
    mu = np.concatenate([np.linspace(-1, 1, 4001), [-1.0, 1.0, 0.0]]).astype(f)
    g32 = f(g)
    got = np.array([O.phase_hg(float(m), float(g32)) for m in mu])
    den = ((f(1) + g32 * g32) - (f(2) * g32) * mu).astype(np.float64)
    num = float(f(0.0795774715459476678844418816863) * (f(1) - g32 * g32))
    rel = np.abs(got / (num * den ** -1.5) - 1)
    assert rel.max() < 6e-7 + 2.0 ** -23, (rel.max(), mu[rel.argmax()])
    # and against the exact phase at the FP32 inputs, where |g| keeps den away from cancellation
    if abs(g) <= 0.6:
        m = mu.astype(np.float64)
        gg = float(g32)
        exact = (1 - gg * gg) / (4 * np.pi * (1 + gg * gg - 2 * gg * m) ** 1.5)
        assert np.abs(got / exact - 1).max() < 2e-6


def test_fresnel_closed_forms():
    n = 1.5
    assert abs(O.fresnel(1.0, 1 / n) - ((n - 1) / (n + 1)) ** 2) < 1e-6   # normal incidence, 0.04
    assert abs(O.fresnel(1.0, n) - ((n - 1) / (n + 1)) ** 2) < 1e-6
    crit = math.sqrt(1 - 1 / n ** 2)  # cos of the critical angle from inside
    assert O.fresnel(crit * 0.999, n) == 1.0          # total internal reflection
    assert O.fresnel(crit * 1.001, n) < 1.0
    assert O.fresnel(1e-6, 1 / n) > 0.999             # grazing
    # Brewster angle: r_p = 0 -> R = r_s^2 / 2
    tb = math.atan(n)
    ci, ct = math.cos(tb), math.sqrt(1 - (math.sin(tb) / n) ** 2)
    rs = ((1 / n) * ci - ct) / ((1 / n) * ci + ct)
    assert abs(O.fresnel(ci, 1 / n) - rs ** 2 / 2) < 1e-5


# ---------------------------------------------------------------- intersections
@pytest.fixture(scope="module")
def room():
    return O.Oracle(O.preset(1), 1)


def test_axis_rays_hit_walls(room):
    o = np.array([[0.5, 0.7, 0.9]] * 5, np.float32)
    d = np.array([[1, 0, 0], [0, -1, 0], [-1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    t, kind, idx = room.intersect(o, d)
    assert kind.tolist() == [1] * 5
    assert idx.tolist() == [0, 1, 2, 3, 4]
    np.testing.assert_allclose(t, [0.5, 0.7, 0.5, 0.3, 0.1], rtol=1e-6)


def test_front_is_open_and_camera_enters(room):
    t, kind, _ = room.intersect(np.array([[0.5, 0.7, 0.9]], np.float32), np.array([[0, 0, -1]], np.float32))
    assert kind[0] == 0
    # From the camera position outside the room, the first hit is inside the room.
    t, kind, idx = room.intersect(np.array([[0.5, 0.5, -0.8]], np.float32), np.array([[0, 0, 1]], np.float32))
    assert kind[0] == 1 and idx[0] == 4 and abs(t[0] - 1.8) < 1e-6


def test_ray_through_sphere_centre(room):
    cfg = O.preset(1)
    for s in range(cfg.n_spheres):
        c = np.array(cfg.spheres[s].center[:], np.float64)
        r = cfg.spheres[s].radius
        o = np.array([c[0], c[1] + 0.02, 0.02])
        d = c - o
        dist = np.linalg.norm(d)
        t, kind, idx = room.intersect(o.astype(np.float32)[None], (d / dist).astype(np.float32)[None])
        assert kind[0] == 2 and idx[0] == s
        assert abs(t[0] - (dist - r)) < 2e-6


# ---------------------------------------------------------------- medium
def test_homogeneous_free_flight_exponential():
    c = O.preset(1)
    o = O.Oracle(c, 1)
    st = c.sigma_s + c.sigma_a
    origin, d = np.array([0.5, 0.5, 0.5], np.float32), np.array([1.0, 0.0, 0.0], np.float32)
    ts = np.array([o.collision(pid, 0, origin, d, 1e9) for pid in range(60000)])
    assert abs(ts.mean() - 1 / st) / (1 / st) < 0.02       # mean free path 1/sigma_t
    for L in (0.5, 2.0, 10.0):
        frac = (ts < L).mean()
        assert abs(frac - (1 - math.exp(-st * L))) < 0.01  # closed-form transmittance


def test_delta_tracking_transmittance_matches_density_integral():
    c = O.preset(2)
    o = O.Oracle(c, 1)
    origin, d = np.array([0.05, 0.5, 0.5], np.float32), np.array([1.0, 0.0, 0.0], np.float32)
    L = 0.9
    n = 40000
    hits = sum(o.collision(pid, 0, origin, d, L) is not None for pid in range(n))
    xs = np.linspace(0, L, 2001)
    pts = np.stack([0.05 + xs, np.full_like(xs, 0.5), np.full_like(xs, 0.5)], 1).astype(np.float32)
    tau = (c.sigma_s + c.sigma_a) * np.trapezoid(o.density(pts).astype(np.float64), xs)
    p_ref = 1 - math.exp(-tau)
    assert abs(hits / n - p_ref) < 4 * math.sqrt(p_ref * (1 - p_ref) / n) + 1e-3


def test_value_noise_grid():
    c = O.preset(2)
    o = O.Oracle(c, 1)
    g = o.noise_grid()  # [z][y][x]
    assert g.shape == (64, 64, 64) and g.min() >= 0.0 and g.max() < 1.0
    # Corner vertices coincide with lattice points: they equal the Philox uniforms.
    lat = c.noise_freq + 1
    def lattice(ix, iy, iz):
        idx = (iz * lat + iy) * lat + ix
        r = O.philox(np.array([[idx, 0, 0, 32]], np.uint32), np.array([[c.seed, 0]], np.uint32))[0, 0]
        return np.float32((r >> 8) * 2.0 ** -24)
    assert g[0, 0, 0] == lattice(0, 0, 0)
    assert g[63, 63, 63] == lattice(8, 8, 8)
    assert g[0, 63, 0] == lattice(0, 8, 0)
    # Density lookups at grid vertices return the vertex value; between vertices
    # the trilinear interpolant stays within the cell's range.
    p = np.array([[i / 63, j / 63, k / 63] for i, j, k in [(3, 5, 7), (10, 20, 30), (62, 1, 0)]], np.float32)
    np.testing.assert_allclose(o.density(p), [g[7, 5, 3], g[30, 20, 10], g[0, 1, 62]], atol=2e-7)
    mid = np.array([[3.5 / 63, 5.5 / 63, 7.5 / 63]], np.float32)
    cell = g[7:9, 5:7, 3:5]
    v = o.density(mid)[0]
    assert cell.min() - 1e-6 <= v <= cell.max() + 1e-6
    np.testing.assert_allclose(v, cell.mean(), atol=1e-6)


def test_presets_match_configs():
    c1, c2, c3 = O.preset(1), O.preset(2), O.preset(3)
    assert (c1.n_emit, c1.k_surface, c1.width, c1.height, c1.spp) == (100000, 32, 256, 256, 1)
    assert abs(c1.rcap_surface - 0.05) < 1e-7 and abs(c1.sigma_s - 0.1) < 1e-7 and abs(c1.sigma_a - 0.01) < 1e-7
    assert (c1.max_bounces, round(c1.rr_survive, 6)) == (8, 0.8)
    assert list(c1.light_pos) == pytest.approx([0.5, 0.98, 0.5]) and c1.light_power[0] == 100.0
    assert (c2.cap_surface, c2.cap_volume, c2.k_volume, c2.width, c2.height, c2.spp) == (1000000, 2000000, 64,
                                                                                        1920, 1080, 4)
    assert c2.medium == pmamd.PM_MEDIUM_NOISE and c2.noise_res == 64 and abs(c2.hg_g - 0.6) < 1e-7
    assert c3.storage == pmamd.PM_STORE_BF16 and c3.fp64_transmittance == 1 and (c3.width, c3.height) == (480, 270)
