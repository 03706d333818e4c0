"""GPU parity: libpm.so (HIP kernels on gfx950, called through the C ABI)
against the CPU oracle on the same seeded inputs.

Bars (from the method definition): photon records, kd-tree node order and
kNN neighbour lists bit-exact; FP64 flux tally within 1e-9 relative; per-pixel
radiance within 1e-4 relative (|gpu - cpu| <= 1e-4 |cpu| + 1e-7).
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle as O
import pmamd

pytestmark = pytest.mark.gpu

RAD_RTOL = 1e-4
RAD_ATOL = 1e-7


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_cfg2(cap_s=60000, cap_v=40000, w=96, h=54, spp=2):
    c = pmamd.preset(2)
    c.cap_surface, c.cap_volume = cap_s, cap_v
    c.width, c.height, c.spp = w, h, spp
    return c


def run_pair(cfg, ncpu, build=True):
    g = pmamd.PhotonMapper(cfg)
    g.emit()
    o = O.Oracle(cfg, ncpu)
    o.emit()
    if build:
        g.build()
        o.build()
    return g, o


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def assert_records_equal(g, o, which):
    gp, gw, gd = g.records(which)
    op, ow, od = o.records(which)
    assert gp.shape == op.shape
    assert np.array_equal(bits(gp), bits(op)), "positions differ"
    if g.cfg.storage == pmamd.PM_STORE_BF16:
        def bf(x):
            u = bits(x.astype(np.float32)).astype(np.uint64)
            r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
            return r.astype(np.uint32)
        assert np.array_equal(bits(gw), bf(ow)), "bf16 power differs"
        assert np.array_equal(bits(gd), bf(od)), "bf16 direction differs"
    else:
        assert np.array_equal(bits(gw), bits(ow)), "power differs"
        assert np.array_equal(bits(gd), bits(od)), "direction differs"


def assert_radiance_close(gpu, cpu):
    err = np.abs(gpu.astype(np.float64) - cpu)
    tol = RAD_RTOL * np.abs(cpu) + RAD_ATOL
    bad = err > tol
    assert not bad.any(), "radiance mismatch at %d px, worst rel %.3g" % (
        bad.sum(), float((err / (np.abs(cpu) + 1e-30)).max()))


@pytest.fixture(scope="module")
def pair1(ncpu):
    say("case: preset 1 (100k photons, homogeneous medium, k=32)")
    return run_pair(pmamd.preset(1), ncpu)


def test_counts_and_tally_preset1(pair1):
    g, o = pair1
    s, c = g.stats(), o.counts()
    assert (s.n_surface, s.n_volume) == (c["n_surface"], c["n_volume"])
    assert (s.n_emit_surface, s.n_emit_volume, s.n_emit_total) == (
        c["n_emit_surface"], c["n_emit_volume"], c["n_emit_total"])
    t = o.tally()
    for name in ("absorbed", "escaped", "rr_gain"):
        np.testing.assert_allclose(np.array(getattr(s, name)[:]), t[name], rtol=1e-9)
    closure = np.array(s.emitted[:]) + np.array(s.rr_gain[:]) - np.array(s.absorbed[:]) - np.array(s.escaped[:])
    assert np.abs(closure).max() < 1e-9 * 100.0


def test_records_bitexact_preset1(pair1):
    g, o = pair1
    assert_records_equal(g, o, pmamd.PM_MAP_SURFACE)
    assert_records_equal(g, o, pmamd.PM_MAP_VOLUME)


def test_tree_bitexact_preset1(pair1):
    g, o = pair1
    for which in (pmamd.PM_MAP_SURFACE, pmamd.PM_MAP_VOLUME):
        gp, gs, ga = g.tree(which)
        op, os_, oa = o.tree(which)
        assert np.array_equal(gp, op), "permutation differs (map %d)" % which
        assert np.array_equal(bits(gs), bits(os_)), "split values differ (map %d)" % which
        assert np.array_equal(ga, oa), "split axes differ (map %d)" % which


def knn_queries(o, which, n_rand=3000, seed=7):
    rng = np.random.default_rng(seed)
    pos = o.records(which)[0]
    q = [rng.random((n_rand, 3)).astype(np.float32)]
    if len(pos):
        sel = rng.integers(0, len(pos), size=min(1000, len(pos)))
        q.append(pos[sel])  # exact hits: d2 == 0 and many near-ties
        q.append((pos[sel] + rng.normal(0, 0.01, (len(sel), 3))).astype(np.float32))
    return np.concatenate(q).astype(np.float32)


@pytest.mark.parametrize("which,k,rcap", [(0, 32, 0.05), (1, 32, 0.05), (0, 64, 0.03), (0, 7, 0.2), (1, 64, 0.5)])
def test_knn_bitexact_preset1(pair1, which, k, rcap):
    g, o = pair1
    q = knn_queries(o, which)
    gi, gd, gc = g.knn(which, q, k, rcap)
    oi, od, oc = o.knn(which, q, k, rcap, brute=True)
    assert np.array_equal(gc, oc)
    assert np.array_equal(gi, oi)
    assert np.array_equal(bits(gd), bits(od))


@pytest.mark.parametrize("mode", [0, 1, 2, 4, 7])
def test_render_crop_preset1(pair1, mode):
    g, o = pair1
    g.set_option(pmamd.PM_OPT_GATHER_MODE, mode)
    x0, y0, w, h = 96, 96, 64, 64
    gimg = g.render(x0, y0, w, h)
    oimg = o.render(x0, y0, w, h)
    assert_radiance_close(gimg, oimg)
    s = g.stats()
    assert (s.surface_queries, s.volume_queries) == o.last_queries


@pytest.mark.parametrize("mode", [0, 1, 2, 4, 7])
def test_render_cone_estimator(ncpu, mode):
    c = pmamd.preset(1)
    c.estimator = pmamd.PM_EST_CONE
    c.n_emit = 30000
    g, o = run_pair(c, ncpu)
    g.set_option(pmamd.PM_OPT_GATHER_MODE, mode)
    assert_radiance_close(g.render(100, 40, 45, 43), o.render(100, 40, 45, 43))


def test_no_medium(ncpu):
    c = pmamd.preset(1)
    c.media_on = 0
    c.n_emit = 30000
    g, o = run_pair(c, ncpu)
    assert g.stats().n_volume == 0
    assert_records_equal(g, o, 0)
    assert_radiance_close(g.render(0, 0, 256, 32), o.render(0, 0, 256, 32))


@pytest.fixture(scope="module")
def pair2(ncpu):
    say("case: preset 2 reduced (caps 60k/40k, value-noise medium, HG 0.6, k=64)")
    return run_pair(small_cfg2(), ncpu)


def test_caps_mode_preset2(pair2):
    g, o = pair2
    s, c = g.stats(), o.counts()
    assert (s.n_surface, s.n_volume) == (60000, 40000)
    assert (s.n_emit_surface, s.n_emit_volume, s.n_emit_total) == (
        c["n_emit_surface"], c["n_emit_volume"], c["n_emit_total"])
    assert_records_equal(g, o, 0)
    assert_records_equal(g, o, 1)


def test_tree_and_knn_preset2(pair2):
    g, o = pair2
    for which in (0, 1):
        gp, gs, ga = g.tree(which)
        op, os_, oa = o.tree(which)
        assert np.array_equal(gp, op) and np.array_equal(bits(gs), bits(os_)) and np.array_equal(ga, oa)
        q = knn_queries(o, which, n_rand=2000)
        gi, gd, gc = g.knn(which, q, 64, 0.03)
        oi, od, oc = o.knn(which, q, 64, 0.03, brute=True)
        assert np.array_equal(gi, oi) and np.array_equal(bits(gd), bits(od)) and np.array_equal(gc, oc)


@pytest.mark.parametrize("mode", [0, 1, 2, 4, 7])
def test_render_preset2_reduced(pair2, mode):
    g, o = pair2
    g.set_option(pmamd.PM_OPT_GATHER_MODE, mode)
    assert_radiance_close(g.render(), o.render())
    s = g.stats()
    assert (s.surface_queries, s.volume_queries) == o.last_queries


@pytest.mark.parametrize("mode", [0, 1, 2, 4, 7])
def test_bf16_preset3(ncpu, mode):
    c = pmamd.preset(3)
    c.n_emit = 30000
    c.width, c.height = 96, 54
    g, o = run_pair(c, ncpu)
    g.set_option(pmamd.PM_OPT_GATHER_MODE, mode)
    assert_records_equal(g, o, 0)
    assert_records_equal(g, o, 1)
    assert_radiance_close(g.render(), o.render(bf16=True))


def test_legacy_abi_matches_pm_api(lib):
    w, h = 64, 64
    buf = pmamd.DeviceBuffer(w * h * 4)
    lib.launch_init_random_numbers_kernel()
    lib.launch_emit_photons_kernel(buf.ptr, w, h, ctypes.c_float(0.0), False, True)
    lib.launch_photon_mapping_kernel(buf.ptr, w, h, ctypes.c_float(0.0), False, True)
    legacy = buf.download(np.uint8, w * h * 4).reshape(h, w, 4)
    c = pmamd.preset(1)
    c.width, c.height = w, h
    g = pmamd.PhotonMapper(c)
    g.emit()
    g.build()
    _, u8 = g.render(want_u8=True)
    assert np.array_equal(legacy, u8)
    assert legacy[..., :3].max() > 0 and (legacy[..., 3] == 0).all()


def test_golden_small_scene(ncpu):
    """libpm.so against the committed oracle fixture (tests/golden/small_scene.npz)."""
    import importlib.util
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    spec = importlib.util.spec_from_file_location("gen_fixture_small", os.path.join(here, "golden",
                                                                                    "gen_fixture_small.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    ref = np.load(os.path.join(here, "golden", "small_scene.npz"))
    g = pmamd.PhotonMapper(mod.small_config())
    g.emit()
    g.build()
    for which, name in ((0, "surface"), (1, "volume")):
        pos, pw, d = g.records(which)
        assert np.array_equal(bits(pos), bits(ref[name + "_pos"]))
        assert np.array_equal(bits(pw), bits(ref[name + "_power"]))
        assert np.array_equal(bits(d), bits(ref[name + "_dir"]))
        perm, split, axis = g.tree(which)
        assert np.array_equal(perm, ref[name + "_perm"]) and np.array_equal(bits(split), bits(ref[name + "_split"]))
        assert np.array_equal(axis, ref[name + "_axis"])
        idx, d2, cnt = g.knn(which, ref["queries"], 16, 0.1)
        assert np.array_equal(idx, ref[name + "_knn_idx"]) and np.array_equal(cnt, ref[name + "_knn_count"])
        assert np.array_equal(bits(d2), bits(ref[name + "_knn_d2"]))
    for mode in (0, 1, 2, 4, 7):
        g.set_option(pmamd.PM_OPT_GATHER_MODE, mode)
        assert_radiance_close(g.render(), ref["render"])


@pytest.mark.parametrize("regions", [1, 2, 4, 8])
def test_tile_regions_same_image(pair1, regions):
    # per-XCD tile regions only change which workgroup renders a tile; the
    # crop is ragged (partial tiles on two sides)
    g, _ = pair1
    g.set_option(pmamd.PM_OPT_GATHER_MODE, 7)
    g.set_option(pmamd.PM_OPT_SAMPLE_LANES, 0)
    try:
        g.set_option(pmamd.PM_OPT_TILE_REGIONS, 0)
        ref = g.render(13, 7, 101, 77)
        g.set_option(pmamd.PM_OPT_TILE_REGIONS, regions)
        img = g.render(13, 7, 101, 77)
    finally:
        g.set_option(pmamd.PM_OPT_TILE_REGIONS, 0)
        g.set_option(pmamd.PM_OPT_SAMPLE_LANES, 1)
    assert np.array_equal(bits(img), bits(ref))


def test_tile_regions_option_checked(pair1):
    g, _ = pair1
    for bad in (-1, 3, 5, 16):
        with pytest.raises(pmamd.PmError):
            g.set_option(pmamd.PM_OPT_TILE_REGIONS, bad)
    for bad in (3, 6, 8):
        with pytest.raises(pmamd.PmError):
            g.set_option(pmamd.PM_OPT_GATHER_MODE, bad)


@pytest.mark.parametrize("spp", [2, 4, 8, 16])
def test_sample_lanes_same_image(pair2, spp):
    # a pixel's samples in adjacent lanes, summed in sample order: the same
    # pixels as one lane looping over the samples (ragged crop)
    g, _ = pair2
    c = g.cfg
    g.set_option(pmamd.PM_OPT_GATHER_MODE, 7)
    try:
        g.set_render(c.width, c.height, spp, c.k_surface, c.k_volume, c.rcap_surface, c.rcap_volume)
        g.set_option(pmamd.PM_OPT_SAMPLE_LANES, 0)
        ref = g.render(5, 3, 61, 37)
        q0 = g.stats()
        g.set_option(pmamd.PM_OPT_SAMPLE_LANES, 1)
        img = g.render(5, 3, 61, 37)
        q1 = g.stats()
    finally:
        g.set_option(pmamd.PM_OPT_SAMPLE_LANES, 1)
        g.set_render(c.width, c.height, c.spp, c.k_surface, c.k_volume, c.rcap_surface, c.rcap_volume)
    assert np.array_equal(bits(img), bits(ref))
    assert (q0.surface_queries, q0.volume_queries) == (q1.surface_queries, q1.volume_queries)


def test_sample_lanes_matches_oracle(pair2):
    g, o = pair2
    g.set_option(pmamd.PM_OPT_GATHER_MODE, 7)
    try:
        g.set_option(pmamd.PM_OPT_SAMPLE_LANES, 0)
        loop = g.render()
        g.set_option(pmamd.PM_OPT_SAMPLE_LANES, 1)
        img = g.render()
    finally:
        g.set_option(pmamd.PM_OPT_SAMPLE_LANES, 1)
    assert np.array_equal(bits(img), bits(loop))
    assert_radiance_close(img, o.render())
    s = g.stats()
    assert (s.surface_queries, s.volume_queries) == o.last_queries
    with pytest.raises(pmamd.PmError):
        g.set_option(pmamd.PM_OPT_SAMPLE_LANES, 2)


@pytest.mark.parametrize("mode", [1, 7])
def test_render_without_march_matches_oracle(pair2, ncpu, mode):
    # PM_OPT_RENDER_MEDIA 0: the same photon maps, radiance estimates at the
    # surface hits only; the oracle renders the GPU's records with the medium
    # switched off
    g, _ = pair2
    s = g.stats()
    c = type(g.cfg).from_buffer_copy(g.cfg)
    c.media_on = 0
    o = O.Oracle(c, ncpu)
    o.set_records(pmamd.PM_MAP_SURFACE, *g.records(pmamd.PM_MAP_SURFACE), s.n_emit_surface)
    o.set_records(pmamd.PM_MAP_VOLUME, *g.records(pmamd.PM_MAP_VOLUME), s.n_emit_volume)
    o.build()
    g.set_option(pmamd.PM_OPT_GATHER_MODE, mode)
    g.set_option(pmamd.PM_OPT_RENDER_MEDIA, 0)
    try:
        img = g.render()
        q = g.stats()
    finally:
        g.set_option(pmamd.PM_OPT_RENDER_MEDIA, 1)
        g.set_option(pmamd.PM_OPT_GATHER_MODE, 5)
    ref = o.render()
    say("surface queries %d, volume queries %d" % (q.surface_queries, q.volume_queries))
    assert q.volume_queries == 0 and q.surface_queries > 0
    assert (q.surface_queries, q.volume_queries) == o.last_queries
    assert_radiance_close(img, ref)
    with pytest.raises(pmamd.PmError):
        g.set_option(pmamd.PM_OPT_RENDER_MEDIA, 2)


def wall_points(rng, n):
    # uniform on the room's five walls (x = 0, x = 1, y = 0, y = 1, z = 1), each
    # arriving along the wall's outward direction
    w = rng.integers(0, 5, n)
    u, v = rng.random(n, dtype=np.float32), rng.random(n, dtype=np.float32)
    pos = np.zeros((n, 3), np.float32)
    dirs = np.zeros((n, 3), np.float32)
    axis = np.array([0, 0, 1, 1, 2])[w]
    side = np.array([0.0, 1.0, 0.0, 1.0, 1.0], np.float32)[w]
    others = np.array([[1, 2], [1, 2], [0, 2], [0, 2], [0, 1]])[w]
    rows = np.arange(n)
    pos[rows, axis] = side
    pos[rows, others[:, 0]] = u
    pos[rows, others[:, 1]] = v
    dirs[rows, axis] = np.where(side > 0.5, 1.0, -1.0)
    return pos, dirs


@pytest.mark.parametrize("storage", ["fp32", "bf16"])
def test_set_records_matches_oracle(ncpu, storage):
    # caller-given surface records (uniform on the walls), rendered without the
    # march, against the oracle on the same records
    c = small_cfg2(w=80, h=48, spp=1)
    c.media_on = 0
    if storage == "bf16":
        c.storage = pmamd.PM_STORE_BF16
    rng = np.random.default_rng(5)
    n = 30000
    pos, dirs = wall_points(rng, n)
    power = rng.uniform(0.5, 1.5, (n, 3)).astype(np.float32)
    g = pmamd.PhotonMapper(c)
    g.set_records(pmamd.PM_MAP_SURFACE, pos, power, dirs, 60000)
    g.build()
    gp, gw, gd = g.records(pmamd.PM_MAP_SURFACE)
    assert np.array_equal(bits(gp), bits(pos))
    if storage == "fp32":
        assert np.array_equal(bits(gw), bits(power)) and np.array_equal(bits(gd), bits(dirs))
    s = g.stats()
    assert (s.n_surface, s.n_emit_surface, s.n_volume) == (n, 60000, 0)
    o = O.Oracle(c, ncpu)
    o.set_records(pmamd.PM_MAP_SURFACE, gp, gw, gd, 60000)
    o.build()
    gp_, gs_, ga_ = g.tree(pmamd.PM_MAP_SURFACE)
    op_, os_, oa_ = o.tree(pmamd.PM_MAP_SURFACE)
    assert np.array_equal(gp_, op_) and np.array_equal(bits(gs_), bits(os_)) and np.array_equal(ga_, oa_)
    img = g.render()
    ref = o.render()
    assert g.stats().surface_queries == o.last_queries[0] > 0
    assert_radiance_close(img, ref)
    with pytest.raises(pmamd.PmError):
        g.set_records(2, pos, power, dirs, 60000)


@pytest.mark.parametrize("estimator", ["box", "cone"])
def test_sample_lanes_bf16_fp64_transmittance(ncpu, estimator):
    # BF16 power/direction storage, FP64 transmittance and both surface
    # estimators with a pixel's 4 samples in adjacent lanes: the same pixels as
    # the one-lane sample loop, and the oracle's within tolerance
    c = pmamd.preset(3)
    c.n_emit = 30000
    c.width, c.height, c.spp = 64, 40, 4
    c.estimator = pmamd.PM_EST_CONE if estimator == "cone" else pmamd.PM_EST_BOX
    g, o = run_pair(c, ncpu)
    g.set_option(pmamd.PM_OPT_GATHER_MODE, 7)
    g.set_option(pmamd.PM_OPT_SAMPLE_LANES, 0)
    loop = g.render()
    g.set_option(pmamd.PM_OPT_SAMPLE_LANES, 1)
    img = g.render()
    assert np.array_equal(bits(img), bits(loop))
    assert_radiance_close(img, o.render(bf16=True))
