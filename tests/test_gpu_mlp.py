"""GPU learned radiance cache (MFMA BF16 kernels through the C ABI) against
the NumPy reference in tests/mlp_ref.py, which is itself pinned by finite
differences and torch autograd (tests/test_mlp_reference.py).

Bars: forward output and every layer's weight/bias gradient within 1e-2
relative (L2 norm) of the reference computed with the same BF16 rounding
points, cosine similarity > 0.9999; training on a smooth target reduces the
loss; repeated calls are bit-identical (deterministic split-K reductions).
"""
import sys

import numpy as np
import pytest

import mlp_ref as R
import pmamd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sample(n, seed=0):
    rng = np.random.default_rng(seed)
    x = np.concatenate([rng.random((n, 3)), rng.normal(size=(n, 3)), np.zeros((n, 3))], 1)
    x[:, 3:6] /= np.linalg.norm(x[:, 3:6], axis=1, keepdims=True)
    x[: n // 2, 6:9] = [0.0, 1.0, 0.0]  # half surface queries (floor normal), half volume
    t = np.exp(rng.normal(size=(n, 3)))
    return x.astype(np.float32), t.astype(np.float32)


@pytest.fixture(scope="module")
def pm():
    c = pmamd.preset(5)
    return pmamd.PhotonMapper(c)


def layer_slices():
    out, p = [], 0
    for o, i in R.SHAPES:
        out.append((slice(p, p + o * i), slice(p + o * i, p + o * i + o)))
        p += o * i + o
    return out


def test_param_layout_and_init(pm):
    assert pmamd.load().pm_cache_param_count() == R.param_count()
    pm.cache_reset(7)
    p = pm.cache_params()
    layers = R.unpack(p.astype(np.float64))
    assert np.all(layers[0][0][:, 216:] == 0) and np.all(layers[4][0][3:] == 0)
    pm.cache_reset(7)
    assert np.array_equal(pm.cache_params(), p)


def test_forward_matches_reference(pm):
    say("case: cache forward, 5000 rows")
    flat = R.init(11)
    pm.cache_params(flat.astype(np.float32))
    x, _ = sample(5000, 1)
    got = pm.cache_predict(x).astype(np.float64)
    ref = R.predict(flat.astype(np.float32).astype(np.float64), x.astype(np.float64), bf=True)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 1e-2, err


@pytest.mark.parametrize("fused_mode", [1, 2, 3, 4])
@pytest.mark.parametrize("n", [1, 37, 5000, 300000, 1100000])
def test_fused_inference_matches_generic(pm, n, fused_mode):
    say("case: fused (mode %d) vs generic inference, %d rows" % (fused_mode, n))
    flat = R.init(13).astype(np.float32)
    pm.cache_params(flat)
    x, _ = sample(n, 5)
    pm.set_option(pmamd.PM_OPT_CACHE_FUSED_INFER, 0)
    gen = pm.cache_predict(x).astype(np.float64)
    pm.set_option(pmamd.PM_OPT_CACHE_FUSED_INFER, fused_mode)
    fused = pm.cache_predict(x).astype(np.float64)
    pm.set_option(pmamd.PM_OPT_CACHE_FUSED_INFER, 4)
    if fused_mode == 1:
        # same K order as the generic GEMMs: differences are FP32 rounding only
        assert np.abs(fused - gen).max() <= 1e-3 * np.abs(gen).max() + 1e-6
    else:
        # K = 32 accumulation order (permuted K in modes 3 and 4), double-angle
        # features and, in mode 4, the normal's constant features of the
        # volume rows summed into the bias change the FP32 rounding ahead of
        # each BF16 conversion, so a few activations land one BF16 step apart;
        # bounded in norm and in the worst case
        assert np.linalg.norm(fused - gen) <= 2e-3 * np.linalg.norm(gen) + 1e-6
        assert np.abs(fused - gen).max() <= 1e-2 * np.abs(gen).max() + 1e-6
    if n <= 5000:
        ref = R.predict(flat.astype(np.float64), x.astype(np.float64), bf=True)
        assert np.linalg.norm(fused - ref) / np.linalg.norm(ref) < 1e-2


@pytest.mark.parametrize("n", [1, 2, 37, 5000, 1100000])
def test_volume_layer1_matches_full(pm, n):
    """Mode 4 sends the surface rows through the full fused kernel (as mode 3,
    bit for bit) and the volume rows through the 5-step layer 1, whose only
    change is the sum of W1's normal columns in the bias."""
    say("case: fused inference, short layer 1 for volume rows, %d rows" % n)
    flat = R.init(17).astype(np.float32)
    pm.cache_params(flat)
    x, _ = sample(n, 9)
    x[n // 2:, 6:9] = 0.0
    pm.set_option(pmamd.PM_OPT_CACHE_FUSED_INFER, 3)
    full = pm.cache_predict(x).astype(np.float64)
    pm.set_option(pmamd.PM_OPT_CACHE_FUSED_INFER, 4)
    short = pm.cache_predict(x).astype(np.float64)
    surf = np.any(x[:, 6:9] != 0.0, axis=1)
    assert np.array_equal(short[surf], full[surf])
    vol = ~surf
    if vol.any():
        d = short[vol] - full[vol]
        say("volume rows: rel L2 %.3g, max %.3g" % (np.linalg.norm(d) / np.linalg.norm(full[vol]),
                                                     np.abs(d).max() / np.abs(full[vol]).max()))
        assert np.linalg.norm(d) <= 1e-3 * np.linalg.norm(full[vol]) + 1e-7
        assert np.abs(d).max() <= 1e-2 * np.abs(full[vol]).max() + 1e-7
    if n <= 5000:
        ref = R.predict(flat.astype(np.float64), x.astype(np.float64), bf=True)
        assert np.linalg.norm(short - ref) / np.linalg.norm(ref) < 1e-2


@pytest.mark.parametrize("n", [256, 4096, 20000])
def test_gradients_match_reference(pm, n):
    say("case: cache gradient, %d rows" % n)
    flat = R.init(12).astype(np.float32)
    pm.cache_params(flat)
    x, t = sample(n, 2)
    g, loss = pm.cache_grad(x, t)
    rl, rg = R.loss_and_grad(flat.astype(np.float64), x.astype(np.float64), t.astype(np.float64), bf=True)
    assert abs(loss - rl) <= 1e-3 * rl
    for l, (ws, bs) in enumerate(layer_slices()):
        for s, name in ((ws, "W"), (bs, "b")):
            a, b = g[s].astype(np.float64), rg[s]
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            cos = a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)
            assert rel < 1e-2 and cos > 0.9999, (l, name, rel, cos)
    g2, _ = pm.cache_grad(x, t)
    assert np.array_equal(g, g2)


def test_training_reduces_loss(pm):
    say("case: cache training, 8 x 16384 rows")
    pm.cache_reset(3)
    rng = np.random.default_rng(4)
    n = 16384 * 8
    x = np.concatenate([rng.random((n, 3)), np.tile([0.0, 0.0, 1.0], (n, 1)), np.zeros((n, 3))], 1).astype(
        np.float32)
    # smooth radiance field over the room
    t = np.stack([2 + np.sin(3 * x[:, 0]), 1 + x[:, 1] ** 2, 3 * np.exp(-x[:, 2])], 1).astype(np.float32)
    first = pm.cache_train(x[:16384], t[:16384], 16384, 1e-3)
    for _ in range(40):
        last = pm.cache_train(x, t, 16384, 1e-3)
    assert last < 0.02 * first, (first, last)
    pred = pm.cache_predict(x[:2000])
    assert np.abs(pred - t[:2000]).mean() / t[:2000].mean() < 0.06
