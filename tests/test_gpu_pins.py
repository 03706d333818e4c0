"""Closed-form pins run on the HIP path (libpm.so through the C ABI): the
march through the k-NN gather, the HG sign of the volume estimate, the uchar4
frame-write rule (pm_render, the cache render and the reference's launchers),
one and two Adam steps of the cache against NumPy, and the error paths
(non-finite configuration, 32-bit record capacity).

The same scenes are pinned on the CPU in test_oracle_pins.py; here the GPU is
compared with the oracle (1e-4 relative per pixel) and with the closed forms.
"""
import ctypes
import math
import sys

import numpy as np
import pytest

import oracle as O
import pmamd
from pin_scenes import hg, load_uniform, march_weights, uniform_ratio, uniform_volume_scene

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def assert_radiance_close(gpu, cpu, rtol=1e-4, atol=1e-7):
    err = np.abs(gpu.astype(np.float64) - cpu)
    bad = err > rtol * np.abs(cpu) + atol
    assert not bad.any(), "radiance mismatch at %d px, worst rel %.3g" % (
        bad.sum(), float((err / (np.abs(cpu) + 1e-30)).max()))


def u8_rule(rgb, exposure):
    """The reference's frame write (photonMappingKernel.cu:1451-1453) in NumPy:
    (unsigned char)min(c * 255.0, 255) in double, with c < 0 and NaN -> 0."""
    c = (rgb.astype(np.float32) * np.float32(exposure)).astype(np.float32).astype(np.float64) * 255.0
    c = np.where(np.isnan(c), 0.0, c)
    return np.floor(np.clip(c, 0.0, 255.0)).astype(np.uint8)


# ------------------------------------------------------------------ uchar4 frame
@pytest.fixture(scope="module")
def small_pair(ncpu):
    say("case: preset 1, 30k photons, 96x64")
    c = pmamd.preset(1)
    c.n_emit = 30000
    c.width, c.height = 96, 64
    g = pmamd.PhotonMapper(c)
    g.emit()
    g.build()
    o = O.Oracle(c, ncpu)
    o.emit()
    o.build()
    return g, o, o.render()


@pytest.mark.parametrize("mode", [0, 1, 7])
def test_u8_frame_follows_reference_rule(small_pair, mode):
    g, o, ref = small_pair
    g.set_option(pmamd.PM_OPT_GATHER_MODE, mode)
    try:
        rgb, u8 = g.render(want_u8=True)
    finally:
        g.set_option(pmamd.PM_OPT_GATHER_MODE, 5)
    assert_radiance_close(rgb, ref)
    e = g.cfg.exposure
    assert np.array_equal(u8[..., :3], u8_rule(rgb, e))
    assert np.array_equal(u8[..., :3], O.tonemap(rgb, e))
    assert (u8[..., 3] == 0).all()
    # against the oracle's FP32 frame: equal except where the 1e-4 radiance
    # tolerance straddles an 8-bit step
    d = np.abs(u8[..., :3].astype(int) - u8_rule(ref, e).astype(int))
    assert d.max() <= 1 and (d == 0).mean() > 0.999
    assert u8[..., :3].max() == 255 and u8[..., :3].min() < 255


def test_u8_cache_frame_follows_reference_rule():
    c = pmamd.preset(5)
    c.n_emit = 30000
    c.width, c.height = 64, 48
    g = pmamd.PhotonMapper(c)
    g.emit()
    g.build()
    g.cache_reset(1)
    g.cache_configure(train_samples=4096, batch=1024)
    rgb, u8 = g.render_cache(want_u8=True)
    assert np.array_equal(u8[..., :3], u8_rule(rgb, c.exposure))
    assert (u8[..., 3] == 0).all()


@pytest.mark.parametrize("interp,anim,media", [(False, 0.0, True), (True, 0.0, True), (False, 1.3, True),
                                               (True, 2.2, False)])
def test_legacy_launchers_against_oracle(lib, ncpu, interp, anim, media):
    # the three launchers in the reference's call order (callbacksPBO.cpp:55-63)
    # against the oracle's frame of the same scene: preset 1 with the
    # animation time, the cone estimate for interpolateFlag and the medium flag
    w, h = 64, 48
    say("case: legacy launchers interpolate=%s anim=%g media=%s" % (interp, anim, media))
    buf = pmamd.DeviceBuffer(w * h * 4)
    lib.launch_init_random_numbers_kernel()
    lib.launch_emit_photons_kernel(buf.ptr, w, h, ctypes.c_float(anim), interp, media)
    lib.launch_photon_mapping_kernel(buf.ptr, w, h, ctypes.c_float(anim), interp, media)
    got = buf.download(np.uint8, w * h * 4).reshape(h, w, 4)
    # the photon maps are reused while no emit call comes (emitFlag off)
    lib.launch_photon_mapping_kernel(buf.ptr, w, h, ctypes.c_float(anim), interp, media)
    again = buf.download(np.uint8, w * h * 4).reshape(h, w, 4)
    assert np.array_equal(got, again)
    c = pmamd.preset(1)
    c.width, c.height, c.spp = w, h, 1
    c.anim_time = anim
    c.estimator = pmamd.PM_EST_CONE if interp else pmamd.PM_EST_BOX
    c.media_on = 1 if media else 0
    o = O.Oracle(c, ncpu)
    o.emit()
    o.build()
    ref = u8_rule(o.render(), c.exposure)
    d = np.abs(got[..., :3].astype(int) - ref.astype(int))
    assert d.max() <= 1 and (d == 0).mean() > 0.999
    assert (got[..., 3] == 0).all()


def test_legacy_animation_moves_spheres(lib):
    # a new animTime re-traces the photons for the moved spheres: the frame
    # changes, and going back to the first time gives the first frame again
    w, h = 64, 48
    buf = pmamd.DeviceBuffer(w * h * 4)
    frames = []
    lib.launch_init_random_numbers_kernel()
    for t in (0.0, 1.0, 0.0):
        lib.launch_emit_photons_kernel(buf.ptr, w, h, ctypes.c_float(t), False, True)
        lib.launch_photon_mapping_kernel(buf.ptr, w, h, ctypes.c_float(t), False, True)
        frames.append(buf.download(np.uint8, w * h * 4))
    assert not np.array_equal(frames[0], frames[1])
    assert np.array_equal(frames[0], frames[2])


# ------------------------------------------------------------------ march and HG through the gather
@pytest.mark.parametrize("mode", [1, 7])
def test_march_through_gather_closed_form(ncpu, mode):
    c, pos, dirs, rho = uniform_volume_scene()
    n_emit = 10000
    say("case: uniform volume photons (1M, k=64), gather mode %d" % mode)
    g = pmamd.PhotonMapper(c)
    load_uniform(g, c, pos, dirs, n_emit)
    g.set_option(pmamd.PM_OPT_GATHER_MODE, mode)
    img = g.render()
    o = O.Oracle(c, ncpu)
    load_uniform(o, c, pos, dirs, n_emit)
    assert_radiance_close(img, o.render())
    r = uniform_ratio(img, o, c, rho, n_emit)
    assert abs(r.mean() - 1) < 0.01, r.mean()
    assert np.all(np.abs(r - 1) < 0.15)


@pytest.mark.parametrize("g_hg", [0.6, -0.3])
def test_hg_sign_closed_form(ncpu, g_hg):
    n_emit = 10000
    res = {}
    for sgn in (1.0, -1.0):
        c, pos, dirs, _ = uniform_volume_scene(n=300_000, g=g_hg, dirs=[0, 0, sgn])
        c.width = c.height = 24
        gm = pmamd.PhotonMapper(c)
        load_uniform(gm, c, pos, dirs, n_emit)
        res[sgn] = gm.render().astype(np.float64)
    o = O.Oracle(c, ncpu)
    for py in range(c.height):
        for px in range(c.width):
            _, rd = march_weights(o, c, px, py)
            u_wo = -float(rd[2])
            want = hg(u_wo, g_hg) / hg(-u_wo, g_hg)
            got = res[1.0][py, px, 0] / res[-1.0][py, px, 0]
            assert abs(got / want - 1) < 1e-4, (px, py, got, want)


# ------------------------------------------------------------------ Adam
def test_adam_steps_match_numpy():
    # two Adam steps (Kingma & Ba 2015: m = b1 m + (1 - b1) g, v = b2 v +
    # (1 - b2) g^2, p -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps),
    # b1 0.9, b2 0.999, eps 1e-8) on the gradients pm_cache_grad returns
    c = pmamd.preset(5)
    g = pmamd.PhotonMapper(c)
    g.cache_reset(3)
    rng = np.random.default_rng(4)
    n, lr = 4096, 1e-3
    # the kernel's constants are the FP32 values of 0.9, 0.999 and 1e-8, and
    # 1 - b1, 1 - b2 are formed in FP32
    f32 = np.float32
    b1, b2, eps = float(f32(0.9)), float(f32(0.999)), float(f32(1e-8))
    one_b1, one_b2 = float(f32(1.0) - f32(0.9)), float(f32(1.0) - f32(0.999))

    def rows():
        x = np.concatenate([rng.random((n, 3)), rng.normal(size=(n, 3)), np.zeros((n, 3))], 1).astype(np.float32)
        x[:, 3:6] /= np.linalg.norm(x[:, 3:6], axis=1, keepdims=True)
        t = rng.uniform(0.0, 5.0, (n, 3)).astype(np.float32)
        return x, t

    p = g.cache_params().astype(np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for step in (1, 2):
        x, t = rows()
        gr, _ = g.cache_grad(x, t)
        gr = gr.astype(np.float64)
        g.cache_train(x, t, n, lr)
        got = g.cache_params().astype(np.float64)
        m = b1 * m + one_b1 * gr
        v = b2 * v + one_b2 * gr * gr
        c1 = float(f32(1.0) - f32(0.9 ** step))
        c2 = float(f32(1.0) - f32(0.999 ** step))
        want = p - lr * (m / c1) / (np.sqrt(v / c2) + eps)
        moved = np.abs(got - p)
        assert moved.max() > 0.5 * lr
        np.testing.assert_allclose(got, want, rtol=4e-7, atol=2e-9)
        # the first step moves every parameter with a nonzero gradient by lr
        if step == 1:
            nz = np.abs(gr) > 1e-4
            np.testing.assert_allclose(moved[nz], lr, rtol=1e-3)
        p = got


# ------------------------------------------------------------------ error paths
def test_nonfinite_config_rejected():
    for mut in (lambda c: c.light_pos.__setitem__(0, float("nan")),
                lambda c: setattr(c, "sigma_s", float("inf")),
                lambda c: c.spheres[0].center.__setitem__(2, float("nan")),
                lambda c: setattr(c, "exposure", float("nan"))):
        c = pmamd.preset(1)
        mut(c)
        with pytest.raises(pmamd.PmError) as e:
            pmamd.PhotonMapper(c)
        assert e.value.status == pmamd.PM_ERR_INVALID


def test_capacity_fixed_count():
    # n_emit * max_bounces records could pass 2^32: refused before any work
    c = pmamd.preset(1)
    c.max_bounces = 64
    c.n_emit = (0xFFFFFFFF // 64) + 1
    g = pmamd.PhotonMapper(c)
    with pytest.raises(pmamd.PmError) as e:
        g.emit()
    assert e.value.status == pmamd.PM_ERR_CAPACITY
    with pytest.raises(pmamd.PmError) as e:
        g.build()
    assert e.value.status == pmamd.PM_ERR_STATE


def test_capacity_unreachable_cap():
    # a thin medium cannot fill 2M volume records within UINT32_MAX / 64
    # photons: emission stops with PM_ERR_CAPACITY instead of wrapping counts
    c = pmamd.preset(2)
    c.max_bounces = 64
    c.sigma_s, c.sigma_a = 1e-5, 0.0
    c.cap_surface, c.cap_volume = 1000, 2000000
    g = pmamd.PhotonMapper(c)
    with pytest.raises(pmamd.PmError) as e:
        g.emit()
    assert e.value.status == pmamd.PM_ERR_CAPACITY


def test_capacity_batch():
    cfgs = [pmamd.batch_scene(1, i) for i in range(2)]
    for c in cfgs:
        c.n_emit = 1 << 27
        c.max_bounces = 8
    with pytest.raises(pmamd.PmError) as e:
        pmamd.Batch(cfgs)
    assert e.value.status == pmamd.PM_ERR_CAPACITY
