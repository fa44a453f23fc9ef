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


def test_bf16_round_known_values():
    # BF16 keeps 8 significand bits: 1 + 2^-7 is representable, 1 + 2^-8 is
    # the tie between 1 and 1 + 2^-7 (to even: 1), 1 + 3 * 2^-8 the tie
    # between 1 + 2^-7 and 1 + 2^-6 (to even: 1 + 2^-6)
    vals = np.array([1.0, 1 + 2**-7, 1 + 2**-8, 1 + 3 * 2**-8, 1 + 2**-8 + 2**-20, -2.5, 3.0e38, 0.0])
    ref = np.array([1.0, 1 + 2**-7, 1.0, 1 + 2**-6, 1 + 2**-7, -2.5, 2.9999799e38, 0.0])
    out = mlp.bf16_round(vals)
    np.testing.assert_array_equal(out[:6], ref[:6])
    assert out[7] == 0.0
    # the largest finite values round within one BF16 step
    assert abs(out[6] - 3.0e38) <= 3.0e38 * 2**-8
    # every rounded value has at most 8 significand bits
    rng = np.random.default_rng(0)
    x = rng.standard_normal(10000) * 10.0 ** rng.integers(-20, 20, 10000)
    y = mlp.bf16_round(x)
    m, _ = np.frexp(y)
    assert np.all(np.abs(m * 2**8 - np.round(m * 2**8)) == 0)
    assert np.all(np.abs(y - x) <= np.abs(x) * 2**-8)


def test_rounded_products_reduce_to_the_exact_form_on_bf16_values():
    # with operands already on the BF16 grid the rounding changes nothing
    rng = np.random.default_rng(3)
    d, h, b = 4, 6, 9
    p = {k: mlp.bf16_round(v + 0.1 * rng.standard_normal(v.shape))
         for k, v in mlp.init_params(d, h, rng).items()}
    x = mlp.bf16_round(rng.standard_normal((b, d)))
    out_r, _ = mlp.forward(p, x, mlp.bf16_round)
    out, (_, z1, h1, z2, h2) = mlp.forward(p, x)
    # the hidden activations are not on the grid: the rounded form differs
    # by at most the BF16 rounding of h1 and h2 propagated through W2, W3
    assert np.max(np.abs(out_r - out)) < 0.05 * np.max(np.abs(out))
    assert np.max(np.abs(out_r - out)) > 0.0
