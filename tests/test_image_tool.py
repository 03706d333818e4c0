"""pm_image: PFM output and the image-comparison metrics.

CPU: `pm_image diff` on synthetic PFM files must report the metrics computed
here independently with numpy (the definitions of the precision report).
GPU: `pm_image render` writes the same FP32 frame as pm_render.
"""
import os
import subprocess

import numpy as np
import pytest

import pmamd

TOOL = os.path.join(pmamd.PKG_ROOT, "pm_image")


def write_pfm(path, img):
    h, w, _ = img.shape
    with open(path, "wb") as f:
        f.write(b"PF\n%d %d\n-1.0\n" % (w, h))
        f.write(np.ascontiguousarray(img[::-1], dtype="<f4").tobytes())


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = (int(v) for v in f.readline().split())
        assert float(f.readline()) < 0
        data = np.frombuffer(f.read(), dtype="<f4")
    return data.reshape(h, w, 3)[::-1]


def metrics(a, ref, e):
    a, ref = a.astype(np.float64), ref.astype(np.float64)
    d = a - ref
    rel = np.abs(d) / np.maximum(np.abs(ref), 1e-30)
    da, db = np.clip(a * e, 0, 1), np.clip(ref * e, 0, 1)
    return {
        "rmse_rel": np.sqrt((d ** 2).mean()) / np.abs(ref).mean(),
        "display_rmse": np.sqrt(((da - db) ** 2).mean()) / db.mean(),
        "median_rel": np.median(rel),
        "p999_rel": np.percentile(rel, 99.9),
        "max_rel": rel[np.abs(ref) > 0.01 * np.abs(ref).mean()].max(),
    }


def test_diff_metrics(tmp_path):
    rng = np.random.default_rng(3)
    ref = rng.gamma(2.0, 3.0, (37, 53, 3)).astype(np.float32)
    ref[5, 7] = 0.0  # a black value: excluded from max_rel, finite in the rest
    a = (ref * (1 + rng.normal(0, 1e-3, ref.shape))).astype(np.float32)
    a[10, 20, 1] *= 1.5  # one large outlier
    pa, pr = str(tmp_path / "a.pfm"), str(tmp_path / "ref.pfm")
    write_pfm(pa, a)
    write_pfm(pr, ref)
    out = subprocess.run([TOOL, "diff", pa, pr, "--exposure", "0.2"], check=True, capture_output=True,
                         text=True).stdout.split()
    got = dict(zip(out[0::2], (float(v) for v in out[1::2])))
    want = metrics(a, ref, 0.2)
    for k, v in want.items():
        assert abs(got[k] - v) <= 1e-5 * abs(v) + 1e-12, (k, got[k], v)
    assert abs(got["max_rel"] - 0.5) < 0.01


def test_diff_rejects_size_mismatch(tmp_path):
    write_pfm(str(tmp_path / "a.pfm"), np.ones((4, 5, 3), np.float32))
    write_pfm(str(tmp_path / "b.pfm"), np.ones((5, 4, 3), np.float32))
    r = subprocess.run([TOOL, "diff", str(tmp_path / "a.pfm"), str(tmp_path / "b.pfm")], capture_output=True)
    assert r.returncode != 0


@pytest.mark.gpu
def test_render_writes_pm_render_frame(tmp_path):
    out = str(tmp_path / "f.pfm")
    print("case: pm_image render preset 1 at 64x48", flush=True)
    subprocess.run([TOOL, "render", "1", out, "--size", "64x48", "--ppm", str(tmp_path / "f.ppm")], check=True)
    c = pmamd.preset(1)
    c.width, c.height = 64, 48
    g = pmamd.PhotonMapper(c)
    g.emit()
    g.build()
    ref = g.render()
    img = read_pfm(out)
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
    assert os.path.getsize(str(tmp_path / "f.ppm")) == len(b"P6\n64 48\n255\n") + 64 * 48 * 3
