"""The reference's frame loop on the HIP path: photon-map reuse while emission
is off (the reference's emitFlag, callbacksPBO.cpp:60-61), checked against
the oracle, and the interactive host pm_view run headless with scripted
keys (callbacksPBO.cpp:104-142)."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
import pmamd

pytestmark = pytest.mark.gpu

PKG = os.path.dirname(pmamd.LIB_PATH)
VIEW = os.path.join(PKG, "pm_view")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def u8_rule(rgb, exposure):
    c = (rgb.astype(np.float32) * np.float32(exposure)).astype(np.float32).astype(np.float64) * 255.0
    c = np.where(np.isnan(c), 0.0, c)
    return np.floor(np.clip(c, 0.0, 255.0)).astype(np.uint8)


def legacy_records(lib, ctx):
    st = pmamd.PmStats()
    pmamd.check(lib.pm_stats_get(ctx, ctypes.byref(st)), "pm_stats_get")
    out = []
    for which, n, n_emit in ((0, st.n_surface, st.n_emit_surface), (1, st.n_volume, st.n_emit_volume)):
        pos = np.empty((n, 3), np.float32)
        pw = np.empty((n, 3), np.float32)
        d = np.empty((n, 3), np.float32)
        pmamd.check(lib.pm_get_records(ctx, which, pos.ctypes.data, pw.ctypes.data, d.ctypes.data, n),
                    "pm_get_records")
        out.append((which, pos, pw, d, n_emit))
    return out


def test_emission_off_reuses_photon_maps(lib, ncpu):
    """Emit at t = 0, then render at t = 1 with no emit call: the spheres the
    eye rays see move, the photon maps stay those of t = 0."""
    w, h = 64, 48
    say("case: legacy frame loop, emission off at a new animTime")
    buf = pmamd.DeviceBuffer(w * h * 4)
    lib.launch_init_random_numbers_kernel()
    lib.launch_emit_photons_kernel(buf.ptr, w, h, ctypes.c_float(0.0), False, True)
    lib.launch_photon_mapping_kernel(buf.ptr, w, h, ctypes.c_float(0.0), False, True)
    at0 = buf.download(np.uint8, w * h * 4).reshape(h, w, 4)
    lib.launch_photon_mapping_kernel(buf.ptr, w, h, ctypes.c_float(1.0), False, True)
    reused = buf.download(np.uint8, w * h * 4).reshape(h, w, 4)
    recs = legacy_records(lib, lib.pm_legacy_context())
    lib.launch_emit_photons_kernel(buf.ptr, w, h, ctypes.c_float(1.0), False, True)
    lib.launch_photon_mapping_kernel(buf.ptr, w, h, ctypes.c_float(1.0), False, True)
    fresh = buf.download(np.uint8, w * h * 4).reshape(h, w, 4)
    assert not np.array_equal(reused, at0)
    assert not np.array_equal(reused, fresh)
    # the oracle: the t = 1 scene with the t = 0 photon maps
    c = pmamd.preset(1)
    c.width, c.height, c.spp = w, h, 1
    c.anim_time = 1.0
    c.media_on = 1
    o = O.Oracle(c, ncpu)
    for which, pos, pw, d, n_emit in recs:
        o.set_records(which, pos, pw, d, n_emit)
    o.build()
    ref = u8_rule(o.render(), c.exposure)
    diff = np.abs(reused[..., :3].astype(int) - ref.astype(int))
    assert diff.max() <= 1 and (diff == 0).mean() > 0.999
    lib.launch_init_random_numbers_kernel()


def run_view(args, timeout=120):
    p = subprocess.run([VIEW, "--headless"] + args, capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, p.stderr
    frames = []
    for line in p.stdout.splitlines():
        m = re.match(r"frame (\d+) t=([\d.]+) emit=(\d) anim=(\d) interp=(\d) media=(\d) voxel=(\d) sum=(\w+)", line)
        if m:
            frames.append(dict(t=float(m.group(2)), emit=int(m.group(3)), anim=int(m.group(4)),
                               interp=int(m.group(5)), media=int(m.group(6)), voxel=int(m.group(7)),
                               sum=m.group(8)))
    return frames


def test_view_headless_keys(tmp_path):
    """pm_view's frame loop with one key per frame: each toggle changes the
    frame, a repeated state repeats it bit for bit, the animation clock
    advances by the step, and Esc ends the loop."""
    say("case: pm_view --headless with scripted keys")
    out = tmp_path / "last.ppm"
    keys = "." + "." + "p" + "." + "a" + "." + "e" + "." + "v" + "i" + "a" + "." + "\x1b" + "."
    f = run_view(["--size", "64x48", "--frames", "14", "--keys", keys, "--out", str(out)])
    assert len(f) == 12  # Esc at frame 12 ends the loop
    assert f[0]["sum"] == f[1]["sum"]                       # same state, same frame
    assert f[2]["media"] == 1 and f[2]["sum"] != f[1]["sum"]  # medium on
    assert f[3]["sum"] == f[2]["sum"]
    assert f[4]["anim"] == 1 and f[4]["t"] == 0.0 and f[4]["sum"] == f[3]["sum"]
    assert abs(f[5]["t"] - 0.1) < 1e-6 and f[5]["sum"] != f[4]["sum"]  # spheres moved, photons re-emitted
    assert f[6]["emit"] == 0 and f[6]["sum"] != f[5]["sum"]  # spheres move, photon maps of t = 0.1 kept
    assert f[8]["voxel"] == 1 and f[8]["sum"] != f[7]["sum"]
    assert f[9]["interp"] == 1 and f[9]["sum"] != f[8]["sum"]  # bilinear window blend
    assert f[10]["anim"] == 0 and f[11]["t"] == f[10]["t"] and f[11]["sum"] == f[10]["sum"]
    data = out.read_bytes()
    assert data.startswith(b"P6\n64 48\n255\n") and len(data) == len(b"P6\n64 48\n255\n") + 64 * 48 * 3
