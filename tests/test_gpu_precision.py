"""Precision variant (the 480x270 BF16 scene): error of the GPU renders against
the FP64 oracle render of the same photon map and eye samples.

The FP64 oracle renders from the FP32 photon records with every per-pixel
operation (camera ray, intersections, march, transmittance, kNN distances,
estimates, sums) in double.  Reported for two GPU forms:
  * plain FP32 (FP32 power/direction, FP32 transmittance);
  * the precision variant: BF16 power and direction, FP32 positions and split
    planes, FP64 transmittance, FP32 compensated accumulation.
Metrics (per pixel and colour channel):
  rmse_rel      sqrt(mean (img - ref)^2) / mean(ref)
  display_rmse  the same on the displayed values (x exposure, clamped to [0,1])
  median_rel    median |img - ref| / ref
  p999_rel      99.9th percentile of |img - ref| / ref
  max_rel       max |img - ref| / ref over values above 1% of mean(ref)
FP32-vs-FP64 differences are concentrated in a few pixels where the two
precisions take different discrete decisions (a silhouette hit or miss, one
march step more or less, a different k-th neighbour); the median shows the
storage cost of BF16.
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


def metrics(img, ref, exposure):
    img = img.astype(np.float64)
    d = img - ref
    rel = np.abs(d) / np.maximum(np.abs(ref), 1e-30)
    big = ref > 0.01 * ref.mean()
    da, db = np.clip(img * exposure, 0, 1), np.clip(ref * exposure, 0, 1)
    return {
        "rmse_rel": float(np.sqrt((d ** 2).mean()) / ref.mean()),
        "display_rmse": float(np.sqrt(((da - db) ** 2).mean()) / db.mean()),
        "median_rel": float(np.median(rel)),
        "p999_rel": float(np.percentile(rel, 99.9)),
        "max_rel": float(rel[big].max()),
    }


@pytest.fixture(scope="module")
def fp64_ref(ncpu):
    c = pmamd.preset(3)
    say("case: FP64 oracle render, %dx%d, %d spp" % (c.width, c.height, c.spp))
    o = O.Oracle(c, ncpu)
    o.emit()
    o.build()
    return c, o.render(fp64=True, bf16=False)


def gpu_render(c):
    g = pmamd.PhotonMapper(c)
    g.emit()
    g.build()
    return g.render()


def test_precision_variant_vs_fp64_oracle(fp64_ref):
    c, ref = fp64_ref
    m_bf = metrics(gpu_render(c), ref, c.exposure)
    c32 = pmamd.preset(3)
    c32.storage, c32.fp64_transmittance = pmamd.PM_STORE_FP32, 0
    m_32 = metrics(gpu_render(c32), ref, c.exposure)
    for name, m in (("BF16 power/dir + FP64 transmittance", m_bf), ("plain FP32", m_32)):
        say("%-36s " % name + " ".join("%s %.3e" % kv for kv in m.items()))
    # FP32 arithmetic: most pixels agree with FP64 to a few ulps
    assert m_32["median_rel"] < 1e-4
    assert m_32["p999_rel"] < 1e-2
    # BF16 storage (8-bit significand, ~2^-9 relative per stored value,
    # averaged over k neighbours) costs well under 1e-3 at the median and
    # nothing visible at the tail
    assert m_bf["median_rel"] < 1e-3
    assert m_bf["p999_rel"] < 1e-2
    assert m_bf["display_rmse"] < 1e-2 and m_32["display_rmse"] < 1e-2
    assert m_bf["rmse_rel"] < 1.5 * m_32["rmse_rel"] + 1e-3
