"""Top-of-tree tiles in LDS (PM_OPT_TOP_LEVELS, pm_batch_tuning.top_levels):
the mode-7 traversals read the top levels of both kd-trees from a copy in
each wave's LDS.  The values are the same, so the frames and query counts
are bit-identical to the render without tiles, for every tile depth, both
sample layouts, the cone estimate and FP64 transmittance."""
import sys

import numpy as np
import pytest

import pmamd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("variant", ["lanes", "loop", "cone", "fp64"])
def test_top_tiles_bit_identical(variant):
    say("case: top-of-tree tiles, preset 2 reduced, %s" % variant)
    c = pmamd.preset(2)
    c.cap_surface, c.cap_volume = 60000, 40000
    c.width, c.height, c.spp = 96, 64, 4
    if variant == "cone":
        c.estimator = pmamd.PM_EST_CONE
    if variant == "fp64":
        c.fp64_transmittance = 1
    g = pmamd.PhotonMapper(c)
    g.emit()
    g.build()
    g.set_option(pmamd.PM_OPT_GATHER_MODE, 7)
    g.set_option(pmamd.PM_OPT_SAMPLE_LANES, 0 if variant == "loop" else 1)
    ref = g.render()
    rs = g.stats()
    for lv in (1, 3, 5, 7):
        g.set_option(pmamd.PM_OPT_TOP_LEVELS, lv)
        img = g.render()
        s = g.stats()
        assert np.array_equal(bits(img), bits(ref)), "top levels %d" % lv
        assert (s.surface_queries, s.volume_queries) == (rs.surface_queries, rs.volume_queries)
    with pytest.raises(pmamd.PmError):
        g.set_option(pmamd.PM_OPT_TOP_LEVELS, 8)
    g.close()


def test_option_ranges():
    """The options added beside the gather modes take only their documented values."""
    c = pmamd.preset(1)
    c.n_emit = 1000
    g = pmamd.PhotonMapper(c)
    for opt, bad in ((pmamd.PM_OPT_TOP_LEVELS, -1), (pmamd.PM_OPT_VOXEL_GRID, 3), (pmamd.PM_OPT_CACHE_NEAR_GROUPS, 17),
                     (pmamd.PM_OPT_CACHE_EMA, 1000), (pmamd.PM_OPT_CACHE_EMA, -1), (pmamd.PM_OPT_CACHE_SEEDS, 2),
                     (99, 0)):
        with pytest.raises(pmamd.PmError):
            g.set_option(opt, bad)
    for opt, ok in ((pmamd.PM_OPT_TOP_LEVELS, 7), (pmamd.PM_OPT_VOXEL_GRID, 2), (pmamd.PM_OPT_CACHE_NEAR_GROUPS, 0),
                    (pmamd.PM_OPT_CACHE_EMA, 0), (pmamd.PM_OPT_CACHE_EMA, 999), (pmamd.PM_OPT_CACHE_SEEDS, 0),
                    (pmamd.PM_OPT_CACHE_SEEDS, 1)):
        g.set_option(opt, ok)
    g.close()
