"""Pins the FP64 surrogate-MLP restatement: gradients by central finite
differences, GELU against its erf-free closed form, Adam against its
published update."""
import numpy as np

from oracle import mlp


def test_gelu_and_derivative():
    z = np.linspace(-6, 6, 1001)
    ref = 0.5 * z * (1 + np.tanh(np.sqrt(2 / np.pi) * (z + 0.044715 * z**3)))
    np.testing.assert_allclose(mlp.gelu(z), ref, rtol=1e-15)
    h = 1e-6
    fd = (mlp.gelu(z + h) - mlp.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(mlp.gelu_grad(z), fd, rtol=1e-7, atol=1e-9)


def test_loss_gradient_finite_differences():
    rng = np.random.default_rng(0)
    d, h, b = 5, 7, 11
    p = mlp.init_params(d, h, rng)
    for k in p:
        p[k] = p[k] + 0.1 * rng.standard_normal(p[k].shape)  # nonzero biases too
    x = rng.standard_normal((b, d))
    t = rng.standard_normal((b, d))
    loss, g = mlp.loss_and_grads(p, x, t)
    eps = 1e-6
    for k in p:
        flat = p[k].ravel()
        for i in rng.choice(flat.size, min(flat.size, 12), replace=False):
            old = flat[i]
            flat[i] = old + eps
            lp, _ = mlp.loss_and_grads(p, x, t)
            flat[i] = old - eps
            lm, _ = mlp.loss_and_grads(p, x, t)
            flat[i] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[k].ravel()[i]) < 1e-7 + 1e-5 * abs(fd), (k, i)


def test_adam_first_step_is_signed_learning_rate():
    p = {"w": np.array([1.0, -2.0, 3.0])}
    g = {"w": np.array([0.5, -0.25, 4.0])}
    m = {"w": np.zeros(3)}
    v = {"w": np.zeros(3)}
    mlp.adam_step(p, g, m, v, 1, lr=0.1)
    # first bias-corrected step: m_hat = g, v_hat = g^2 -> update = lr * g / (|g| + eps)
    np.testing.assert_allclose(p["w"], [1.0 - 0.1, -2.0 + 0.1, 3.0 - 0.1], rtol=1e-7)
