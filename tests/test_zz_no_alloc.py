"""pm_render_crop_noalloc: the render with its query counters in the
library's per-GPU scratch array instead of an allocation.  Called through the
C ABI, it gives the same image bits, uchar4 frame and query counts as
pm_render_crop on the same context, on two calls in a row (the second on a
smaller crop, so the counters it leaves behind would show if not cleared)."""
import sys

import numpy as np
import pytest

import pmamd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def render_with(lib, fn, g, x0, y0, w, h):
    rgb = pmamd.DeviceBuffer(w * h * 12)
    u8 = pmamd.DeviceBuffer(w * h * 4)
    pmamd.check(fn(g.h, x0, y0, w, h, rgb.ptr, u8.ptr), fn.__name__)
    s = g.stats()
    return (rgb.download(np.float32, w * h * 3).view(np.uint32), u8.download(np.uint8, w * h * 4),
            (s.surface_queries, s.volume_queries))


@pytest.mark.parametrize("mode", [5, 1, 7])
def test_noalloc_render_matches_render(lib, mode):
    say("case: pm_render_crop_noalloc vs pm_render_crop, gather mode %d" % mode)
    c = pmamd.preset(2)
    c.cap_surface, c.cap_volume = 60000, 40000
    c.width, c.height, c.spp = 160, 96, 2
    g = pmamd.PhotonMapper(c)
    g.emit()
    g.build()
    g.set_option(pmamd.PM_OPT_GATHER_MODE, mode)
    for (x0, y0, w, h) in ((0, 0, 160, 96), (17, 9, 61, 43)):
        ref = render_with(lib, lib.pm_render_crop, g, x0, y0, w, h)
        got = render_with(lib, lib.pm_render_crop_noalloc, g, x0, y0, w, h)
        assert np.array_equal(got[0], ref[0])
        assert np.array_equal(got[1], ref[1])
        assert got[2] == ref[2] and got[2][1] > 0
    g.close()


def test_noalloc_render_argument_checks(lib):
    c = pmamd.preset(1)
    c.n_emit = 5000
    c.width, c.height = 32, 32
    g = pmamd.PhotonMapper(c)
    buf = pmamd.DeviceBuffer(32 * 32 * 12)
    # before the maps exist: call-order error; outside the frame: invalid
    assert lib.pm_render_crop_noalloc(g.h, 0, 0, 32, 32, buf.ptr, None) == pmamd.PM_ERR_STATE
    g.emit()
    g.build()
    assert lib.pm_render_crop_noalloc(g.h, 0, 0, 33, 32, buf.ptr, None) == pmamd.PM_ERR_INVALID
    assert lib.pm_render_crop_noalloc(None, 0, 0, 32, 32, buf.ptr, None) == pmamd.PM_ERR_INVALID
    assert lib.pm_render_crop_noalloc(g.h, 0, 0, 32, 32, buf.ptr, None) == pmamd.PM_OK
