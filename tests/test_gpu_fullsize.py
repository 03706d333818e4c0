"""Parity at the sizes the benchmark reports.

* The 1080p workload (1M surface + 2M volume records, value-noise medium, HG
  g = 0.6, k = 64, r <= 0.03): both maps' records bit-exact against the
  oracle; both kd-trees (permutation, split values, axes) bit-exact; kNN lists
  of a few thousand queries bit-exact against the oracle's tree search; a
  64x64 crop of the 1920x1080 x 4 spp frame within 1e-4 relative per pixel.
* The 64-scene batch (the generator's scenes at full size): every scene's
  records, flux tally, trees and image bit-identical to the scene run alone
  with the same leaf size and gather, and the query counts add up.
"""
import sys

import numpy as np
import pytest

import oracle as O
import pmamd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def full2(ncpu):
    c = pmamd.preset(2)
    say("case: preset 2 at full size (%d / %d records), oracle on %d threads" % (c.cap_surface, c.cap_volume, ncpu))
    g = pmamd.PhotonMapper(c)
    g.emit()
    g.build()
    o = O.Oracle(c, ncpu)
    o.emit()
    o.build()
    return g, o


def test_full_counts_and_records(full2):
    g, o = full2
    s, c = g.stats(), o.counts()
    assert (s.n_surface, s.n_volume) == (1000000, 2000000)
    assert (s.n_emit_surface, s.n_emit_volume, s.n_emit_total) == (
        c["n_emit_surface"], c["n_emit_volume"], c["n_emit_total"])
    for which in (pmamd.PM_MAP_SURFACE, pmamd.PM_MAP_VOLUME):
        for a, b in zip(g.records(which), o.records(which)):
            assert np.array_equal(bits(a), bits(b)), "map %d records differ" % which


def test_full_trees(full2):
    g, o = full2
    for which in (pmamd.PM_MAP_SURFACE, pmamd.PM_MAP_VOLUME):
        gp, gs, ga = g.tree(which)
        op, os_, oa = o.tree(which)
        assert np.array_equal(gp, op), "permutation differs (map %d)" % which
        assert np.array_equal(bits(gs), bits(os_)), "split values differ (map %d)" % which
        assert np.array_equal(ga, oa), "split axes differ (map %d)" % which


@pytest.mark.parametrize("which", [0, 1])
def test_full_knn_lists(full2, which):
    g, o = full2
    rng = np.random.default_rng(21 + which)
    pos = o.records(which)[0]
    sel = rng.integers(0, len(pos), 1500)
    q = np.concatenate([rng.random((2000, 3)), pos[sel],
                        pos[sel] + rng.normal(0, 0.005, (len(sel), 3))]).astype(np.float32)
    gi, gd, gc = g.knn(which, q, 64, 0.03)
    oi, od, oc = o.knn(which, q, 64, 0.03)
    assert np.array_equal(gc, oc)
    assert np.array_equal(gi, oi)
    assert np.array_equal(bits(gd), bits(od))
    assert (gc == 64).mean() > 0.5


@pytest.mark.parametrize("x0,y0,mode", [(928, 508, 7), (300, 700, 7), (1400, 200, 5)])
def test_full_frame_crop(full2, x0, y0, mode):
    # gather mode 7 (the full frame's automatic choice) and the automatic
    # choice for a crop this small, on crops of the full frame
    g, o = full2
    g.set_option(pmamd.PM_OPT_GATHER_MODE, mode)
    try:
        img = g.render(x0, y0, 64, 64)
    finally:
        g.set_option(pmamd.PM_OPT_GATHER_MODE, 5)
    s = g.stats()
    ref = o.render(x0, y0, 64, 64)
    assert (s.surface_queries, s.volume_queries) == o.last_queries
    err = np.abs(img.astype(np.float64) - ref)
    bad = err > 1e-4 * np.abs(ref) + 1e-7
    say("crop (%d, %d): %d volume queries, worst rel %.3g" % (x0, y0, s.volume_queries,
                                                              float((err / (np.abs(ref) + 1e-30)).max())))
    assert not bad.any()


def test_batch64_equals_scenes_alone():
    cfgs = [pmamd.batch_scene(1, i) for i in range(64)]
    say("case: the 64-scene batch at full size (%d photons)" % sum(c.n_emit for c in cfgs))
    b = pmamd.Batch(cfgs)
    b.set_tuning(32, 64, 32, 8, 0)
    b.submit()
    st = b.stats()
    tot_q = 0
    for i, c in enumerate(cfgs):
        c.leaf_size = 32
        g = pmamd.PhotonMapper(c)
        g.set_option(pmamd.PM_OPT_GATHER_MODE, 1)  # the one-scene gather with 8-byte heap keys
        g.emit()
        g.build()
        img = g.render()
        s1, s2 = g.stats(), b.scene_stats(i)
        assert (s1.n_surface, s1.n_volume) == (s2.n_surface, s2.n_volume), i
        for name in ("absorbed", "escaped", "rr_gain"):
            assert list(getattr(s1, name)) == list(getattr(s2, name)), (i, name)
        for which in (pmamd.PM_MAP_SURFACE, pmamd.PM_MAP_VOLUME):
            for a, bb in zip(g.records(which), b.records(i, which)):
                assert np.array_equal(bits(a), bits(bb)), (i, which)
            gp, gs, ga = g.tree(which)
            bp, bs, ba = b.tree(i, which, 32)
            assert np.array_equal(gp, bp) and np.array_equal(bits(gs), bits(bs)) and np.array_equal(ga, ba), (i, which)
        assert np.array_equal(bits(img), bits(b.image(i))), "scene %d image differs" % i
        tot_q += s1.surface_queries + s1.volume_queries
        g.close()
    assert st.surface_queries + st.volume_queries == tot_q
    b.close()
