"""NumPy FP64 reference of the learned radiance cache (test infrastructure).

Same network and parameter layout as libpm.so's pm_cache_* entry points:
9 inputs -> sin/cos(2^b pi v), b = 0..11 (216 features, padded to 224) ->
4 x (128, ReLU) -> 16 outputs (first 3 used) predicting log(1 + radiance);
loss = mean over rows and the 3 channels of (y - log(1 + t))^2.

With bf16=True the values the GPU stores in BF16 are rounded the same way
(features, weights, hidden activations, back-propagated deltas); sums are
FP64 here and FP32 on the GPU.
"""
import numpy as np

IN, BANDS, FEAT, KPAD, HID, OUT, OUTPAD = 9, 12, 216, 224, 128, 3, 16
SHAPES = [(HID, KPAD), (HID, HID), (HID, HID), (HID, HID), (OUTPAD, HID)]


def param_count():
    return sum(o * i + o for o, i in SHAPES)


def unpack(flat):
    out, p = [], 0
    for o, i in SHAPES:
        W = flat[p:p + o * i].reshape(o, i)
        p += o * i
        b = flat[p:p + o]
        p += o
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def bf16(x):
    x32 = np.asarray(x, np.float32)
    u = x32.view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def encode(x):
    x = np.asarray(x, np.float64).reshape(-1, IN)
    n = x.shape[0]
    f = np.zeros((n, KPAD))
    for i in range(IN):
        for b in range(BANDS):
            a = 2 * np.pi * x[:, i] * 2.0 ** (b - 1)
            f[:, i * 2 * BANDS + 2 * b] = np.sin(a)
            f[:, i * 2 * BANDS + 2 * b + 1] = np.cos(a)
    return f


def forward(flat, x, bf=True):
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    layers = unpack(np.asarray(flat, np.float64))
    acts = [r(encode(x))]
    for l in range(4):
        W, b = layers[l]
        z = acts[-1] @ r(W).T + b
        acts.append(r(np.maximum(z, 0.0)))
    W, b = layers[4]
    y = acts[-1] @ r(W).T + b
    if bf:
        y = y.astype(np.float32).astype(np.float64)
    return acts, y


def loss_and_grad(flat, x, t, bf=True):
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    layers = unpack(np.asarray(flat, np.float64))
    acts, y = forward(flat, x, bf)
    n = y.shape[0]
    tt = np.log1p(np.maximum(np.asarray(t, np.float64).reshape(-1, OUT), 0.0))
    d = y[:, :OUT] - tt
    loss = (d ** 2).mean()
    dy = np.zeros((n, OUTPAD))
    dy[:, :OUT] = 2.0 * d / (n * OUT)
    delta = r(dy)
    grads = [None] * 5
    for l in range(4, -1, -1):
        W, _ = layers[l]
        grads[l] = (delta.T @ acts[l], delta.sum(0))
        if l == 0:
            break
        delta = r((delta @ r(W)) * (acts[l] > 0))
    return loss, pack(grads)


def predict(flat, x, bf=True):
    _, y = forward(flat, x, bf)
    return np.maximum(np.expm1(y[:, :OUT]), 0.0)


def init(seed=1):
    rng = np.random.default_rng(seed)
    layers = []
    for l, (o, i) in enumerate(SHAPES):
        ir = FEAT if l == 0 else HID
        orr = OUT if l == 4 else HID
        s = np.sqrt(6.0 / (ir + orr)) if l == 4 else np.sqrt(6.0 / ir)
        W = np.zeros((o, i))
        W[:orr, :ir] = rng.uniform(-s, s, (orr, ir))
        b = np.zeros(o)
        b[:orr] = rng.uniform(-0.1, 0.1, orr)
        layers.append((W, b))
    return pack(layers)
