"""The batched workload's scene generator (host code of libpm.so, no GPU):
distributions and constraints it promises, determinism, and that the oracle
traces its scenes (all three materials, up to 256 spheres) with flux closure.
"""
import math

import numpy as np

import oracle as O
import pmamd


def scenes(seed=1, n=64):
    return [pmamd.batch_scene(seed, i) for i in range(n)]


def test_generator_ranges_and_constraints():
    cs = scenes()
    mats = []
    for c in cs:
        assert 2 <= c.n_spheres <= 256
        assert 10_000 <= c.n_emit <= 2_000_000
        assert c.width == c.height and c.width in (128, 256, 512, 1024)
        assert 0.0 <= c.sigma_s <= 0.3 and abs(c.sigma_a - 0.1 * c.sigma_s) <= 1e-7
        assert (c.k_surface, c.k_volume, c.spp) == (32, 32, 1)
        lp = np.array(c.light_pos[:])
        ctr = np.array([c.spheres[j].center[:] for j in range(c.n_spheres)])
        rad = np.array([c.spheres[j].radius for j in range(c.n_spheres)])
        assert ((rad >= 0.02) & (rad <= 0.2)).all()
        assert ((ctr >= rad[:, None]) & (ctr <= 1 - rad[:, None])).all(), "sphere leaves the room"
        assert (np.linalg.norm(ctr - lp, axis=1) > rad).all(), "light inside a sphere"
        d = np.linalg.norm(ctr[:, None] - ctr[None], axis=2) + np.eye(len(rad)) * 10
        assert (d > rad[:, None] + rad[None]).all(), "overlapping spheres"
        for j in range(c.n_spheres):
            s = c.spheres[j]
            mats.append(s.material)
            if s.material == pmamd.PM_MAT_DIELECTRIC:
                assert abs(s.ior - 1.5) < 1e-7
    mats = np.array(mats)
    n = len(mats)
    assert n > 2000
    for m, p in ((pmamd.PM_MAT_DIFFUSE, 0.6), (pmamd.PM_MAT_MIRROR, 0.25), (pmamd.PM_MAT_DIELECTRIC, 0.15)):
        # binomial 5-sigma band
        assert abs((mats == m).mean() - p) < 5 * math.sqrt(p * (1 - p) / n)
    # photon counts log-uniform on [1e4, 2e6]: mean of log10 = (4 + 6.301) / 2, sd of the mean
    # (6.301 - 4) / sqrt(12 * 64)
    lg = np.log10([c.n_emit for c in cs])
    assert abs(lg.mean() - 5.1505) < 5 * 2.301 / math.sqrt(12 * len(cs))
    assert len({c.width for c in cs}) == 4


def test_generator_deterministic_and_distinct():
    a, b = scenes(3, 8), scenes(3, 8)
    for x, y in zip(a, b):
        assert bytes(x) == bytes(y)
    assert len({bytes(x) for x in a}) == len(a)
    assert bytes(scenes(4, 1)[0]) != bytes(a[0])


def test_oracle_traces_generated_scene_with_flux_closure():
    c = max(scenes(1, 16), key=lambda c: c.n_spheres)  # the most cluttered of the first 16
    c.n_emit = 20000
    o = O.Oracle(c, 4)
    o.emit()
    cnt, t = o.counts(), o.tally()
    assert cnt["n_surface"] > 0
    closure = np.array(c.light_power[:]) + t["rr_gain"] - t["absorbed"] - t["escaped"]
    assert np.abs(closure).max() < 1e-9 * 100.0
