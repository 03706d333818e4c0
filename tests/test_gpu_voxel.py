"""The voxel-grid estimate (PM_OPT_VOXEL_GRID, the reference's 32^3 grid
method) on the HIP path, through the C ABI, against the oracle: the
summed-volume tables bit for bit (integer work), the frames bit for bit
(the same FP32 transmittance and FP64 window arithmetic on both sides, the
new-records case within 1e-5 relative per pixel), the reference's launchers with the option set on the legacy
context, and pm_render_crop_noalloc in this mode.  The oracle's window
estimates are pinned in test_oracle_voxel.py."""
import ctypes
import sys

import numpy as np
import pytest

import oracle as O
import pmamd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def assert_radiance_close(gpu, cpu, rtol=1e-5, atol=1e-8):
    err = np.abs(gpu.astype(np.float64) - cpu)
    bad = err > rtol * np.abs(cpu) + atol
    assert not bad.any(), "radiance mismatch at %d px, worst rel %.3g" % (
        bad.sum(), float((err / (np.abs(cpu) + 1e-30)).max()))


def u8_rule(rgb, exposure):
    c = (rgb.astype(np.float32) * np.float32(exposure)).astype(np.float32).astype(np.float64) * 255.0
    c = np.where(np.isnan(c), 0.0, c)
    return np.floor(np.clip(c, 0.0, 255.0)).astype(np.uint8)


def oracle_with_records(c, g, ncpu):
    """An oracle holding the GPU context's records (the emission's parity is
    checked elsewhere)."""
    o = O.Oracle(c, ncpu)
    s = g.stats()
    for which, n_emit in ((pmamd.PM_MAP_SURFACE, s.n_emit_surface), (pmamd.PM_MAP_VOLUME, s.n_emit_volume)):
        pos, pw, d = g.records(which)
        o.set_records(which, pos, pw, d, n_emit)
    return o


CASES = [
    # preset, storage, media, photons, (w, h, spp)
    (1, pmamd.PM_STORE_FP32, 1, 40000, (96, 64, 2)),
    (1, pmamd.PM_STORE_BF16, 1, 40000, (64, 48, 1)),
    (2, pmamd.PM_STORE_FP32, 1, 60000, (80, 48, 1)),
    (1, pmamd.PM_STORE_FP32, 0, 30000, (64, 48, 1)),
]


@pytest.mark.parametrize("case", CASES, ids=["p1-fp32", "p1-bf16", "p2-density", "p1-nomedia"])
def test_voxel_render_against_oracle(case, ncpu):
    preset, storage, media, n, (w, h, spp) = case
    say("case: voxel grid preset %d storage %d media %d photons %d %dx%dx%d" % (preset, storage, media, n, w, h,
                                                                                 spp))
    c = pmamd.preset(preset)
    c.n_emit = n
    c.cap_surface = c.cap_volume = 0
    c.storage = storage
    c.media_on = media
    c.width, c.height, c.spp = w, h, spp
    g = pmamd.PhotonMapper(c)
    g.emit()  # the voxel estimate needs no kd-tree
    o = oracle_with_records(c, g, ncpu)
    bf16 = storage == pmamd.PM_STORE_BF16
    _, _, ts, tv = o.voxel_tables(bf16=bf16)
    for mode in (1, 2):
        g.set_option(pmamd.PM_OPT_VOXEL_GRID, mode)
        rgb, u8 = g.render(want_u8=True)
        if mode == 1:
            assert np.array_equal(g.voxel_table(pmamd.PM_MAP_SURFACE), ts)
            assert np.array_equal(g.voxel_table(pmamd.PM_MAP_VOLUME), tv)
        ref = o.render_voxel(interp=mode == 2, bf16=bf16)
        s = g.stats()
        say("  mode %d: %.4f of the pixels bit-exact, %d surface / %d volume queries" % (
            mode, float((rgb.astype(np.float64) == ref).all(-1).mean()), s.surface_queries, s.volume_queries))
        # the window sums are exact integers and the arithmetic around them
        # is the same IEEE sequence on both sides: the frames agree bit for bit
        assert np.array_equal(rgb.astype(np.float64), ref)
        assert (s.surface_queries, s.volume_queries) == o.last_queries
        assert np.array_equal(u8[..., :3], u8_rule(rgb, c.exposure))
    # the k-NN render still needs pm_build; with it, the option off gives the k-NN frame
    g.set_option(pmamd.PM_OPT_VOXEL_GRID, 0)
    with pytest.raises(pmamd.PmError):
        g.render()
    g.close()


def test_voxel_grids_follow_new_records(ncpu):
    """pm_set_records after a voxel render: the next render rebuilds the grids."""
    say("case: voxel grids rebuilt after pm_set_records")
    c = pmamd.preset(1)
    c.n_emit = 20000
    c.width, c.height, c.spp = 48, 32, 1
    g = pmamd.PhotonMapper(c)
    g.emit()
    g.set_option(pmamd.PM_OPT_VOXEL_GRID, 1)
    first = g.render()
    rng = np.random.default_rng(3)
    pos = rng.random((5000, 3), dtype=np.float32)
    pw = np.full((5000, 3), 0.5, np.float32)
    d = np.tile(np.float32([0, 0, 1]), (5000, 1))
    g.set_records(pmamd.PM_MAP_VOLUME, pos, pw, d, 5000)
    second = g.render()
    o = oracle_with_records(c, g, ncpu)
    assert np.array_equal(g.voxel_table(pmamd.PM_MAP_VOLUME), o.voxel_tables()[3])
    assert_radiance_close(second, o.render_voxel())
    assert not np.array_equal(first, second)
    g.close()


def test_voxel_noalloc_matches_render(lib):
    say("case: pm_render_crop_noalloc in voxel-grid mode")
    c = pmamd.preset(1)
    c.n_emit = 20000
    c.width, c.height, c.spp = 64, 48, 1
    g = pmamd.PhotonMapper(c)
    g.emit()
    g.set_option(pmamd.PM_OPT_VOXEL_GRID, 2)
    rgb = pmamd.DeviceBuffer(64 * 48 * 12)
    # the grids are built by an allocating render first
    assert lib.pm_render_crop_noalloc(g.h, 0, 0, 64, 48, rgb.ptr, None) == pmamd.PM_ERR_STATE
    for (x0, y0, w, h) in ((0, 0, 64, 48), (5, 7, 30, 20)):
        pmamd.check(lib.pm_render_crop(g.h, x0, y0, w, h, rgb.ptr, None), "pm_render_crop")
        ref = rgb.download(np.float32, w * h * 3).view(np.uint32).copy()
        rq = (g.stats().surface_queries, g.stats().volume_queries)
        pmamd.check(lib.pm_render_crop_noalloc(g.h, x0, y0, w, h, rgb.ptr, None), "pm_render_crop_noalloc")
        got = rgb.download(np.float32, w * h * 3).view(np.uint32)
        assert np.array_equal(got, ref)
        assert (g.stats().surface_queries, g.stats().volume_queries) == rq
    g.close()


@pytest.mark.parametrize("interp", [False, True])
def test_legacy_launchers_voxel_grid(lib, ncpu, interp):
    """The reference's launchers with PM_OPT_VOXEL_GRID set on the legacy
    context: interpolateFlag picks the bilinear blend."""
    w, h = 64, 48
    say("case: legacy launchers, voxel grid, interpolate=%s" % interp)
    buf = pmamd.DeviceBuffer(w * h * 4)
    lib.launch_init_random_numbers_kernel()
    ctx = lib.pm_legacy_context()
    assert ctx
    try:
        pmamd.check(lib.pm_set_option(ctx, pmamd.PM_OPT_VOXEL_GRID, 1), "pm_set_option")
        lib.launch_emit_photons_kernel(buf.ptr, w, h, ctypes.c_float(0.0), interp, True)
        lib.launch_photon_mapping_kernel(buf.ptr, w, h, ctypes.c_float(0.0), interp, True)
        got = buf.download(np.uint8, w * h * 4).reshape(h, w, 4)
        st = pmamd.PmStats()
        pmamd.check(lib.pm_stats_get(ctx, ctypes.byref(st)), "pm_stats_get")
        c = pmamd.preset(1)
        c.width, c.height, c.spp = w, h, 1
        c.estimator = pmamd.PM_EST_CONE if interp else pmamd.PM_EST_BOX
        o = O.Oracle(c, ncpu)
        for which, n, n_emit in ((0, st.n_surface, st.n_emit_surface), (1, st.n_volume, st.n_emit_volume)):
            pos = np.empty((n, 3), np.float32)
            pw = np.empty((n, 3), np.float32)
            d = np.empty((n, 3), np.float32)
            pmamd.check(lib.pm_get_records(ctx, which, pos.ctypes.data, pw.ctypes.data, d.ctypes.data, n),
                        "pm_get_records")
            o.set_records(which, pos, pw, d, n_emit)
        ref = u8_rule(o.render_voxel(interp=interp), c.exposure)
        diff = np.abs(got[..., :3].astype(int) - ref.astype(int))
        assert diff.max() <= 1 and (diff == 0).mean() > 0.999
        assert (got[..., 3] == 0).all()
    finally:
        # a fresh legacy context (the k-NN estimate) for the tests after this one
        lib.launch_init_random_numbers_kernel()
