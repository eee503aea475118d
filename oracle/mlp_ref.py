"""CPU reference of the treelet-proposal MLP. TEST INFRASTRUCTURE ONLY.

FP64 forward / backward of the 102 -> 128 -> 128 -> 128 -> 4096 network in the flat parameter
layout of include/gprt_hip.h (W1t b1 W2t b2 W3t b3 W4t b4, Wt[out][in], first-layer input padded
to 128). `bf16=True` rounds weights, inputs and hidden activations to bf16 (round-to-nearest-even)
at the points where the device stores them, so the forward matches the device up to FP32
accumulation order. The backward is the textbook softmax-cross-entropy / ReLU chain rule, pinned
by central finite differences in tests/test_oracle_mlp.py; Adam follows Kingma & Ba (2015),
bias-corrected.
"""
from __future__ import annotations

import numpy as np

BANDS = 8
IN = 6 + 6 * 2 * BANDS  # 102
IN_PAD = 128
HID = 128
OUT = 4096
W = HID * IN_PAD
W4 = OUT * HID
PARAMS = 3 * (W + HID) + W4 + OUT


def offsets():
    """[(weight offset, weight shape, bias offset, bias size)] for the 4 layers."""
    out = []
    for layer in range(3):
        w0 = layer * (W + HID)
        out.append((w0, (HID, IN_PAD), w0 + W, HID))
    w0 = 3 * (W + HID)
    out.append((w0, (OUT, HID), w0 + W4, OUT))
    return out


def unpack(params: np.ndarray):
    """-> [(Wt [out, in], b [out])] views of the flat vector."""
    res = []
    for w0, shape, b0, bn in offsets():
        res.append((params[w0:w0 + shape[0] * shape[1]].reshape(shape), params[b0:b0 + bn]))
    return res


def bf16_round(x) -> np.ndarray:
    """Round to bf16 (RNE) and return as float64 (NaN is kept)."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isnan(f), np.nan, out)


def encode(rays: np.ndarray, bf16: bool = True) -> np.ndarray:
    """rays: RAY_DTYPE records -> [n, 128] features: o, d, sin/cos(pi 2^k v) per component."""
    v = np.stack([rays["ox"], rays["oy"], rays["oz"], rays["dx"], rays["dy"], rays["dz"]], 1).astype(np.float64)
    n = len(v)
    X = np.zeros((n, IN_PAD))
    X[:, :6] = v
    for c in range(6):
        for k in range(BANDS):
            a = np.pi * (2.0 ** k) * v[:, c]
            X[:, 6 + c * 2 * BANDS + 2 * k] = np.sin(a)
            X[:, 6 + c * 2 * BANDS + 2 * k + 1] = np.cos(a)
    return bf16_round(X) if bf16 else X


def forward(params: np.ndarray, X: np.ndarray, bf16: bool = True):
    """Returns (logits [n, 4096], activations [X, H1, H2, H3], pre-activations [Z1, Z2, Z3])."""
    layers = unpack(np.asarray(params, dtype=np.float64))
    rnd = bf16_round if bf16 else (lambda a: a)
    acts, pres = [X], []
    h = X
    for Wt, b in layers[:3]:
        z = h @ rnd(Wt).T + b
        pres.append(z)
        h = rnd(np.maximum(z, 0.0))
        acts.append(h)
    Wt, b = layers[3]
    logits = h @ rnd(Wt).T + b
    return logits, acts, pres


def loss_and_grad(params: np.ndarray, X: np.ndarray, labels: np.ndarray, bf16: bool = True,
                  bf16_grads: bool = False, with_cond: bool = False):
    """Mean softmax cross-entropy over the rows and its gradient in the flat layout (FP64).

    bf16_grads: also round the backpropagated quantities where the device stores them in bf16
    (the output-layer gradient dZ after scaling by 1/n, and each hidden layer's masked dA), so the
    result differs from the device's only by FP32 accumulation order and the rare bf16 rounding
    flips that order causes. with_cond: also return the elementwise condition sums of the same
    products with every term taken in absolute value (|dZ|^T |H| for a weight gradient, column sums
    of |dZ| for a bias gradient): an accumulation-order or rounding error of relative size e in
    the terms changes an element by at most e times its condition sum."""
    params = np.asarray(params, dtype=np.float64)
    logits, acts, pres = forward(params, X, bf16)
    n = len(X)
    zmax = logits.max(1, keepdims=True)
    e = np.exp(logits - zmax)
    p = e / e.sum(1, keepdims=True)
    loss = float(np.mean(np.log(e.sum(1)) + zmax[:, 0] - logits[np.arange(n), labels]))
    dz = p.copy()
    dz[np.arange(n), labels] -= 1.0
    dz /= n
    layers = unpack(params)
    rnd = bf16_round if bf16 else (lambda a: a)
    grnd = bf16_round if bf16_grads else (lambda a: a)
    grad = np.zeros(PARAMS)
    cond = np.zeros(PARAMS) if with_cond else None
    offs = offsets()
    # output layer
    dz = grnd(dz)
    w0, shape, b0, bn = offs[3]
    grad[w0:w0 + shape[0] * shape[1]] = (dz.T @ acts[3]).ravel()
    grad[b0:b0 + bn] = dz.sum(0)
    if with_cond:
        cond[w0:w0 + shape[0] * shape[1]] = (np.abs(dz).T @ np.abs(acts[3])).ravel()
        cond[b0:b0 + bn] = np.abs(dz).sum(0)
    da = grnd((dz @ rnd(layers[3][0])) * (pres[2] > 0))
    for layer in (2, 1, 0):
        w0, shape, b0, bn = offs[layer]
        grad[w0:w0 + shape[0] * shape[1]] = (da.T @ acts[layer]).ravel()
        grad[b0:b0 + bn] = da.sum(0)
        if with_cond:
            cond[w0:w0 + shape[0] * shape[1]] = (np.abs(da).T @ np.abs(acts[layer])).ravel()
            cond[b0:b0 + bn] = np.abs(da).sum(0)
        if layer > 0:
            da = grnd((da @ rnd(layers[layer][0])) * (pres[layer - 1] > 0))
    return (loss, grad, cond) if with_cond else (loss, grad)


def logit_cond(params: np.ndarray, acts) -> np.ndarray:
    """Elementwise condition sums of the output layer, |H3| |W4|^T + |b4| (bf16 weights): an
    error of relative size e in its terms changes a logit by at most e times this."""
    W4, b4 = unpack(np.asarray(params, dtype=np.float64))[3]
    return np.abs(acts[3]) @ np.abs(bf16_round(W4)).T + np.abs(b4)


def adam_step(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def he_init(seed: int) -> np.ndarray:
    """A random parameter vector of the right shape (not the device's initialiser)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS)
    for layer, (w0, shape, b0, bn) in enumerate(offsets()):
        fan = IN if layer == 0 else HID
        Wt = rng.uniform(-1, 1, shape) * np.sqrt(6.0 / fan)
        if layer == 0:
            Wt[:, IN:] = 0.0
        p[w0:w0 + shape[0] * shape[1]] = Wt.ravel()
        p[b0:b0 + bn] = rng.uniform(-0.1, 0.1, bn)
    return p
