"""Cache render (record -> train -> MLP -> fallback -> accumulate) on the GPU.

* With the fallback threshold at 0 every query is gathered exactly, so the
  pipeline must reproduce the plain render (1e-4 relative, as the plain
  parity bar): this pins recording, weights and accumulation.
* After online training, the cache render of the cache scene is compared to a
  64-spp plain render and must be closer to it than a plain render with the
  same number of eye samples (what the learned part gains).
"""
import sys

import numpy as np
import pytest

import pmamd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def scene():
    c = pmamd.preset(5)
    c.width, c.height = 96, 72
    g = pmamd.PhotonMapper(c)
    g.emit()
    g.build()
    return g


def rel_rmse(a, b, exposure):
    # on the displayed values (exposure, clamp to [0, 1]): the few unclamped
    # pixels beside the point light would otherwise dominate the metric
    da, db = np.clip(a * exposure, 0, 1), np.clip(b * exposure, 0, 1)
    return float(np.sqrt(((da - db) ** 2).mean()) / db.mean())


def test_all_fallback_equals_plain(scene):
    say("case: cache render with threshold 0 (all exact)")
    g = scene
    g.cache_configure(threshold=0.0, train_samples=16384, batch=16384)
    img = g.render_cache()
    ref = g.render()
    s = g.cache_stats()
    assert s.fallbacks == s.records > 0
    err = np.abs(img.astype(np.float64) - ref)
    assert (err <= 1e-4 * np.abs(ref) + 1e-6).all(), float(err.max())


def test_cache_beats_equal_sample_plain(scene):
    say("case: cache render after online training vs 64-spp plain")
    g = scene
    c = g.cfg
    g.cache_reset(1)
    g.cache_configure(threshold=0.05, train_samples=1 << 16, batch=1 << 14)
    for _ in range(12):
        img = g.render_cache()
    s = g.cache_stats()
    say("records %d fallbacks %d loss %.4g" % (s.records, s.fallbacks, s.train_loss))
    assert s.fallbacks < s.records
    plain1 = g.render()
    g.set_render(c.width, c.height, 64, c.k_surface, c.k_volume, c.rcap_surface, c.rcap_volume)
    ref = g.render()
    g.set_render(c.width, c.height, 1, c.k_surface, c.k_volume, c.rcap_surface, c.rcap_volume)
    e_cache, e_plain = rel_rmse(img, ref, c.exposure), rel_rmse(plain1, ref, c.exposure)
    say("display rel RMSE vs 64 spp: cache %.4f, plain 1 spp %.4f; fallback fraction %.3f"
        % (e_cache, e_plain, s.fallbacks / s.records))
    assert e_cache < e_plain
