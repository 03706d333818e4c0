"""Batched scenes on the GPU: one hipGraph launch renders every scene.

Bars:
  * every scene's photon records, flux tally, kd-trees and image are
    bit-identical to the same scene run alone (pm_create / pm_emit /
    pm_build / pm_render with the same leaf size);
  * workgroup size and heap slots never change a bit; another leaf size
    gives the same trees as the scene's own build with that leaf size and the
    same neighbour sets, summed in another order (agreement to 1e-5);
  * one batched scene agrees with the CPU oracle within the plain parity bar
    (1e-4 relative per pixel).
The scenes come from the batched workload's generator (random spheres of all
three materials, homogeneous media of random density), shrunk to test sizes.
"""
import sys

import numpy as np
import pytest

import oracle as O
import pmamd

pytestmark = pytest.mark.gpu

SIZES = [(40, 40, 9000), (64, 48, 20000), (72, 72, 14000), (96, 64, 31000), (56, 56, 6000), (80, 80, 26000)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def small_scenes(seed=7):
    cfgs = []
    for i, (w, h, n) in enumerate(SIZES):
        c = pmamd.batch_scene(seed, i)
        c.width, c.height, c.n_emit = w, h, n
        cfgs.append(c)
    cfgs[2].medium = pmamd.PM_MEDIUM_NONE  # one scene without a medium: empty volume map
    return cfgs


@pytest.fixture(scope="module")
def batch():
    cfgs = small_scenes()
    say("case: batch of %d scenes, spheres %s" % (len(cfgs), [c.n_spheres for c in cfgs]))
    b = pmamd.Batch(cfgs)
    b.submit()
    return cfgs, b


MODE_OF_KEY_BYTES = {8: 1, 4: 2, 2: 4}  # gather mode of the one-scene render with the same heap keys


def run_alone(c, leaf, key_bytes=2, gather_mode=0):
    c.leaf_size = leaf
    g = pmamd.PhotonMapper(c)
    g.set_option(pmamd.PM_OPT_GATHER_MODE, 7 if gather_mode == 7 else MODE_OF_KEY_BYTES[key_bytes])
    g.emit()
    g.build()
    return g, g.render()


def test_batch_equals_scenes_run_alone(batch):
    cfgs, b = batch
    st = b.stats()
    say("graph nodes %d, work items %d, tuning %s" % (st.graph_nodes, st.work_items,
                                                      (st.tuning.leaf_size, st.tuning.block_threads,
                                                       st.tuning.heap_slots, st.tuning.heap_key_bytes,
                                                       st.tuning.gather_mode)))
    L = st.tuning.leaf_size
    tot_q = 0
    for i, c in enumerate(cfgs):
        g, img = run_alone(c, L, st.tuning.heap_key_bytes, st.tuning.gather_mode)
        s1, s2 = g.stats(), b.scene_stats(i)
        assert (s1.n_surface, s1.n_volume) == (s2.n_surface, s2.n_volume), i
        for name in ("absorbed", "escaped", "rr_gain"):
            assert list(getattr(s1, name)) == list(getattr(s2, name)), (i, name)
        for which in (pmamd.PM_MAP_SURFACE, pmamd.PM_MAP_VOLUME):
            for a, bb in zip(g.records(which), b.records(i, which)):
                assert np.array_equal(bits(a), bits(bb)), (i, which)
            gp, gs, ga = g.tree(which)
            bp, bs, ba = b.tree(i, which, L)
            assert np.array_equal(gp, bp) and np.array_equal(bits(gs), bits(bs)) and np.array_equal(ga, ba), (i, which)
        assert np.array_equal(bits(img), bits(b.image(i))), "scene %d image differs" % i
        tot_q += s1.surface_queries + s1.volume_queries
    assert st.surface_queries + st.volume_queries == tot_q


def test_batch_time_phases_same_frame(batch):
    # the frame launched kernel by kernel leaves the graph frame's images and
    # records, and times three phases that fit inside one frame
    cfgs, b = batch
    b.submit()
    imgs = [b.image(i) for i in range(len(cfgs))]
    recs = [b.records(i, pmamd.PM_MAP_SURFACE)[0] for i in range(len(cfgs))]
    ms = b.time_phases()
    say("phases (emit, build, render) ms: %s" % (ms,))
    assert all(t > 0.0 for t in ms)
    for i in range(len(cfgs)):
        assert np.array_equal(bits(imgs[i]), bits(b.image(i))), i
        assert np.array_equal(bits(recs[i]), bits(b.records(i, pmamd.PM_MAP_SURFACE)[0])), i
    b.submit()  # the graph still runs after the eager frame
    for i in range(len(cfgs)):
        assert np.array_equal(bits(imgs[i]), bits(b.image(i))), i


def assert_same_radiance(ref, img, bitwise, what):
    if bitwise:
        assert np.array_equal(bits(ref), bits(img)), what
    else:
        # another leaf size visits the same neighbours in another order: the
        # compensated sums agree to rounding
        np.testing.assert_allclose(img, ref, rtol=1e-5, atol=1e-7, err_msg=what)


@pytest.mark.parametrize("tuning", [(32, 256, 32, 2, 0, 0), (32, 128, 64, 8, 0, 0), (4, 256, 32, 4, 0, 0),
                                    (16, 192, 64, 2, 0, 0), (32, 64, 64, 2, 7, 0), (8, 256, 64, 2, 7, 0),
                                    (32, 128, 64, 2, 7, 3), (16, 256, 64, 2, 7, 5), (32, 64, 64, 2, 7, 7)])
def test_batch_tuning_invariance(batch, tuning):
    cfgs, b = batch
    b.set_tuning(32, 64, 32, 2)
    b.submit()
    ref = [b.image(i) for i in range(len(cfgs))]
    b.set_tuning(*tuning)
    b.submit()
    for i in range(len(cfgs)):
        assert_same_radiance(ref[i], b.image(i), tuning[0] == 32, "scene %d under tuning %s" % (i, tuning))
    # with the same leaf size and gather the batch matches the scene's own run bit for bit
    g, img = run_alone(cfgs[3], tuning[0], tuning[3], tuning[4])
    assert np.array_equal(bits(img), bits(b.image(3)))
    gp, gs, ga = g.tree(pmamd.PM_MAP_SURFACE)
    bp, bs, ba = b.tree(3, pmamd.PM_MAP_SURFACE, tuning[0])
    assert np.array_equal(gp, bp) and np.array_equal(bits(gs), bits(bs)) and np.array_equal(ga, ba)


def test_batch_autotune(batch):
    cfgs, b = batch
    b.set_tuning(32, 64, 32, 2)
    b.submit()
    ref = [b.image(i) for i in range(len(cfgs))]
    best, ms = b.autotune(sample_stride=3)
    say("autotune: leaf %d, block %d, heap slots %d, key bytes %d, gather mode %d, top levels %d "
        "(%.2f ms on the sampled frame)" % (best + (ms,)))
    assert best[0] in (4, 8, 16, 32) and best[1] in (64, 128, 256) and best[2] in (32, 64) and ms > 0
    assert best[3] in (8, 4, 2) and best[4] in (0, 7) and best[5] in (0, 3, 5, 7)
    assert best[4] == 0 or (best[3], best[2]) == (2, 64)
    assert best[5] == 0 or best[4] == 7
    b.submit()
    for i in range(len(cfgs)):
        assert_same_radiance(ref[i], b.image(i), best[0] == 32, "scene %d after autotune" % i)


def test_batch_scene_vs_oracle(batch, ncpu):
    cfgs, b = batch
    b.set_tuning(32, 64, 32, 2)
    b.submit()
    for i in (0, 3):
        c = cfgs[i]
        c.leaf_size = 32
        o = O.Oracle(c, ncpu)
        o.emit()
        o.build()
        cpu = o.render()
        gpu = b.image(i).astype(np.float64)
        err = np.abs(gpu - cpu)
        bad = err > 1e-4 * np.abs(cpu) + 1e-7
        say("scene %d (%d spheres): worst rel %.3g" % (i, c.n_spheres, float((err / (np.abs(cpu) + 1e-30)).max())))
        assert not bad.any(), "scene %d: %d pixels off" % (i, int(bad.sum()))
