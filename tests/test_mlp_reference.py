"""Pins of the NumPy MLP reference used to check the GPU cache kernels (CPU):
gradients against FP64 finite differences and against torch's CPU autograd
on identical weights; the encoding against its closed form."""
import numpy as np
import pytest

import mlp_ref as R


def sample(n, seed=0):
    rng = np.random.default_rng(seed)
    x = np.concatenate([rng.random((n, 3)), rng.normal(size=(n, 3)), rng.normal(size=(n, 3))], 1)
    x[:, 3:6] /= np.linalg.norm(x[:, 3:6], axis=1, keepdims=True)
    t = np.exp(rng.normal(size=(n, 3)))
    return x, t


def test_encoding_closed_form():
    x = np.array([[0.25, 0.5, 0.125, 1.0, -1.0, 0.0, 0.3, 0.0, 0.0]])
    f = R.encode(x)
    assert f.shape == (1, 224) and np.all(f[:, 216:] == 0)
    # input 0 = 0.25: band b has angle 2 pi 0.25 2^(b-1) -> b=0: pi/4, b=1: pi/2, b=2: pi
    np.testing.assert_allclose(f[0, 0:6], [np.sin(np.pi / 4), np.cos(np.pi / 4), 1, 0, 0, -1], atol=1e-12)


def test_gradient_matches_finite_differences():
    flat = R.init(3)
    x, t = sample(16, 1)
    _, g = R.loss_and_grad(flat, x, t, bf=False)
    rng = np.random.default_rng(5)
    idx = rng.choice(np.nonzero(g)[0], 40, replace=False)
    for i in idx:
        h = 1e-6 * max(1.0, abs(flat[i]))
        fp, fm = flat.copy(), flat.copy()
        fp[i] += h
        fm[i] -= h
        num = (R.loss_and_grad(fp, x, t, bf=False)[0] - R.loss_and_grad(fm, x, t, bf=False)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9


def test_gradient_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    flat = R.init(4)
    x, t = sample(64, 2)
    loss, g = R.loss_and_grad(flat, x, t, bf=False)
    params = [torch.tensor(a, dtype=torch.float64, requires_grad=True) for W, b in R.unpack(flat) for a in (W, b)]
    h = torch.tensor(R.encode(x), dtype=torch.float64)
    for l in range(4):
        h = torch.relu(h @ params[2 * l].T + params[2 * l + 1])
    y = h @ params[8].T + params[9]
    tt = torch.log1p(torch.tensor(t, dtype=torch.float64))
    tl = ((y[:, :3] - tt) ** 2).mean()
    tl.backward()
    tg = np.concatenate([p.grad.numpy().ravel() for p in params])
    assert abs(tl.item() - loss) < 1e-12
    np.testing.assert_allclose(g, tg, rtol=1e-9, atol=1e-14)


def test_bf16_rounding():
    # BF16 keeps 8 significant bits: ulp(1) = 2^-7; ties go to the even mantissa.
    v = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -9, -3.0e-5, 65504.0], np.float64)
    r = R.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                    # tie -> even (1.0000000)
    assert r[2] == 1 + 2 ** -6            # tie -> even (1.0000010), not 1 + 2^-7
    assert r[3] == 1.0
    assert np.all(np.abs(r - v) <= np.abs(v) * 2 ** -8)
