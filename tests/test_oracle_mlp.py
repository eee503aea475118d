"""Pins the CPU MLP reference (no GPU): central finite differences of the FP64 loss against the
analytic gradient, the Adam closed form, and the bf16 rounding rule."""
import numpy as np

from gprt_amd import scenes
from oracle import mlp_ref as M


def test_bf16_rounding_rule():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0], dtype=np.float32)
    r = M.bf16_round(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0); 1 + 3*2^-8 rounds up.
    np.testing.assert_array_equal(r, [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0])


def test_encoding_layout():
    rays = scenes.random_rays(4, seed=1)
    X = M.encode(rays, bf16=False)
    assert X.shape == (4, 128)
    np.testing.assert_allclose(X[:, 0], rays["ox"])
    np.testing.assert_allclose(X[:, 6 + 2 * 3], np.sin(np.pi * 8 * rays["ox"].astype(np.float64)))
    np.testing.assert_allclose(X[:, 6 + 16 * 5 + 2 * 7 + 1], np.cos(np.pi * 128 * rays["dz"].astype(np.float64)))
    assert np.all(X[:, M.IN:] == 0)


def test_gradient_finite_differences():
    p = M.he_init(3)
    rays = scenes.random_rays(12, seed=2)
    X = M.encode(rays, bf16=False)
    labels = np.random.default_rng(0).integers(0, M.OUT, len(X))
    loss, g = M.loss_and_grad(p, X, labels, bf16=False)
    assert np.isfinite(loss)
    rng = np.random.default_rng(5)
    idx = []
    for w0, shape, b0, bn in M.offsets():
        idx += list(w0 + rng.integers(0, min(shape[0] * shape[1], 102 * shape[0]), 4)) + list(b0 + rng.integers(0, bn, 2))
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        num = (M.loss_and_grad(pp, X, labels, bf16=False)[0] - M.loss_and_grad(pm, X, labels, bf16=False)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 + 1e-4 * abs(g[i]), (i, num, g[i])


def test_adam_first_step_closed_form():
    w = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -0.01, 4.0])
    w1, m, v = M.adam_step(w, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    # first bias-corrected step moves every weight by lr * sign(g) (up to eps)
    np.testing.assert_allclose(w1, w - 0.01 * np.sign(g), rtol=0, atol=1e-7)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_bf16_gradient_emulation_and_condition_sums():
    """The bf16-rounded backward (the device's storage points) differs from the unrounded FP64
    backward by a small but non-zero share of every tensor, and the condition sums bound
    |gradient| and |logits|."""
    p = M.he_init(4)
    rays = scenes.random_rays(64, seed=3)
    X = M.encode(rays)
    labels = np.random.default_rng(2).integers(0, M.OUT, len(X))
    _, exact = M.loss_and_grad(p, X, labels, bf16=True)
    _, g, cond = M.loss_and_grad(p, X, labels, bf16=True, bf16_grads=True, with_cond=True)
    assert np.all(np.abs(g) <= cond * (1 + 1e-12) + 1e-300)
    # Rounding dZ and dA changes every gradient tensor, by about 0.1-0.3% of its norm.
    for w0, shape, b0, bn in M.offsets():
        for sl in (slice(w0, w0 + shape[0] * shape[1]), slice(b0, b0 + bn)):
            rel = np.linalg.norm(g[sl] - exact[sl]) / np.linalg.norm(exact[sl])
            assert 1e-4 < rel < 1e-2, rel
    z, acts, _ = M.forward(p, X, bf16=True)
    assert np.all(np.abs(z) <= M.logit_cond(p, acts) * (1 + 1e-12))
