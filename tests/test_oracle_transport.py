"""Transport, estimators and render of the oracle, pinned by conservation laws,
closed forms and statistics (CPU)."""
import math
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def run(cfg, nt=8):
    o = O.Oracle(cfg, nt)
    o.emit()
    return o


@pytest.mark.parametrize("preset,caps", [(1, None), (2, (40000, 30000)), (1, "nomedia")])
def test_flux_tally_closes(preset, caps):
    c = O.preset(preset)
    if preset == 1:
        c.n_emit = 50000
    if caps == "nomedia":
        c.media_on = 0
    elif caps:
        c.cap_surface, c.cap_volume = caps
    o = run(c)
    t = o.tally()
    closure = t["emitted"] + t["rr_gain"] - t["absorbed"] - t["escaped"]
    assert np.abs(closure).max() < 1e-10 * t["emitted"].max()
    assert (t["absorbed"] > 0).all() and (t["escaped"] > 0).all()


def test_russian_roulette_is_unbiased():
    # Expected power arriving at surfaces per emitted photon is the same with
    # and without roulette (bounce cap 8 in both).
    def arriving(p):
        c = O.preset(1)
        c.media_on = 0
        c.n_emit = 200000
        c.rr_survive = p
        o = run(c)
        pos, pw, _ = o.records(0)
        return pw.astype(np.float64).sum(0) / c.n_emit, len(pos)

    a08, n08 = arriving(0.8)
    a10, n10 = arriving(1.0)
    assert n10 > n08
    np.testing.assert_allclose(a08, a10, rtol=0.02)


def test_records_lie_on_surfaces_and_direct_hits_come_from_light():
    c = O.preset(1)
    c.n_emit = 20000
    c.media_on = 0
    o = run(c)
    pos, pw, d = o.records(0)
    sph = [(np.array(c.spheres[i].center[:]), c.spheres[i].radius) for i in range(c.n_spheres)]
    dist = np.minimum.reduce([np.abs(pos[:, 0]), np.abs(pos[:, 0] - 1), np.abs(pos[:, 1]), np.abs(pos[:, 1] - 1),
                              np.abs(pos[:, 2] - 1)] +
                             [np.abs(np.linalg.norm(pos - cc, axis=1) - r) for cc, r in sph])
    assert dist.max() < 2e-5
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    direct = np.all(pw == 1.0, axis=1)  # no bounce yet: emitted straight from the light
    assert direct.sum() > 0.5 * c.n_emit
    light = np.array(c.light_pos[:], np.float64)
    v = pos[direct] - light
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    # the records with mirror history also have power 1; most direct hits match the light direction
    match = np.abs((v * d[direct]).sum(1) - 1) < 1e-5
    assert match.mean() > 0.9


def lattice_disc(spacing, radius):
    g = np.arange(-radius, radius + 1e-9, spacing)
    xx, zz = np.meshgrid(g, g)
    keep = xx ** 2 + zz ** 2 <= radius ** 2
    return np.stack([0.5 + xx[keep], np.zeros(keep.sum()), 0.5 + zz[keep]], 1).astype(np.float32)


def test_surface_estimate_uniform_disc_closed_form():
    # Photons on a square lattice of spacing a on the floor: irradiance
    # E = (Phi / N_emit) / a^2 per photon, L = albedo / pi * E.
    a, n_emit = 0.004, 10000
    pos = lattice_disc(a, 0.2)
    n = len(pos)
    c = O.preset(1)
    c.k_surface, c.rcap_surface = 64, 0.1
    o = O.Oracle(c, 2)
    dirs = np.tile(np.array([0, -1, 0], np.float32), (n, 1))
    o.set_records(0, pos, np.ones((n, 3), np.float32), dirs, n_emit)
    o.set_records(1, pos[:0], pos[:0], pos[:0], n_emit)
    o.build()
    albedo = np.array([0.7, 0.5, 0.3])
    E = 100.0 / n_emit / a ** 2
    for fp64 in (False, True):
        L = o.gather(0, [0.5, 0.0, 0.5], [0, 1, 0], albedo, fp64=fp64)
        np.testing.assert_allclose(L, albedo / math.pi * E, rtol=0.05)
    # photons arriving from behind the surface do not count
    L = o.gather(0, [0.5, 0.0, 0.5], [0, -1, 0], albedo)
    assert np.all(L == 0)
    # fewer than k photons inside the cap: r = r_cap, estimate = sum / (pi r_cap^2)
    c2 = O.preset(1)
    c2.k_surface, c2.rcap_surface = 64, 0.01
    o2 = O.Oracle(c2, 2)
    o2.set_records(0, pos, np.ones((n, 3), np.float32), dirs, n_emit)
    o2.set_records(1, pos[:0], pos[:0], pos[:0], n_emit)
    o2.build()
    inside = (((pos - np.array([0.5, 0, 0.5], np.float32)) ** 2).sum(1) <= np.float32(0.01) ** 2).sum()
    assert inside < 64
    L = o2.gather(0, [0.5, 0.0, 0.5], [0, 1, 0], albedo, fp64=True)
    np.testing.assert_allclose(L, albedo / math.pi * (100.0 / n_emit) * inside / (math.pi * 0.01 ** 2), rtol=1e-6)


def test_volume_estimate_poisson_closed_form():
    # Photons uniform at random (density rho) with the isotropic phase: the
    # k-NN volume estimate k / V_k has expectation rho k / (k - 1) for a
    # Poisson process, so the in-scatter estimate averages to
    # (Phi / N_emit) * (1 / 4 pi) * rho * k / (k - 1).
    n, n_emit, k = 200000, 5000, 64
    rng = np.random.default_rng(1)
    pos = (0.2 + 0.6 * rng.random((n, 3))).astype(np.float32)
    rho = n / 0.6 ** 3
    c = O.preset(1)
    c.k_volume, c.rcap_volume, c.hg_g = k, 0.2, 0.0
    o = O.Oracle(c, 2)
    dirs = rng.normal(size=(n, 3)).astype(np.float32)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    o.set_records(1, pos, np.ones((n, 3), np.float32), dirs, n_emit)
    o.set_records(0, pos[:0], pos[:0], pos[:0], n_emit)
    o.build()
    q = (0.35 + 0.3 * rng.random((400, 3))).astype(np.float32)
    S = np.mean([o.gather(1, x, [0, 0, 1], fp64=True)[0] for x in q])
    want = (100.0 / n_emit) / (4 * math.pi) * rho * k / (k - 1)
    assert abs(S / want - 1) < 0.015


def test_render_deterministic_and_fp32_close_to_fp64():
    c = O.preset(1)
    c.n_emit = 20000
    c.width, c.height = 48, 32
    o = run(c)
    o.build()
    a = o.render()
    b = o.render()
    assert np.array_equal(a, b)
    d = o.render(fp64=True)
    rel = np.abs(a - d).sum() / np.abs(d).sum()
    assert rel < 5e-3
    assert (a >= 0).all() and a.mean() > 0.1


def test_golden_fixture_reproduced():
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen_fixture_small",
                                                  os.path.join(HERE, "golden", "gen_fixture_small.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    ref = np.load(os.path.join(HERE, "golden", "small_scene.npz"))
    got = mod.generate()
    for k in ref.files:
        if ref[k].dtype == np.float32:
            assert np.array_equal(ref[k].view(np.uint32), got[k].view(np.uint32)), k
        else:
            assert np.array_equal(ref[k], got[k]), k
