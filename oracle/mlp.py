"""FP64 numpy restatement of the gradient-surrogate MLP and its training step.

TEST INFRASTRUCTURE ONLY (the checker of the HIP training and inference
kernels).  The reference has no learned component; this network follows the
specification of the surrogate-gradient HMC configuration: D -> H -> H -> D,
GELU (tanh form) after the two hidden layers, mean-squared error on
standardised targets, Adam.

Parameters: dict with W1 (H x D), b1 (H), W2 (H x H), b2 (H), W3 (D x H), b3 (D).
"""
import numpy as np

C_GELU = np.sqrt(2.0 / np.pi)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + 0.044715 * z**3)))


def gelu_grad(z):
    u = C_GELU * (z + 0.044715 * z**3)
    t = np.tanh(u)
    du = C_GELU * (1.0 + 3 * 0.044715 * z**2)
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * du


def forward(p, x):
    z1 = x @ p["W1"].T + p["b1"]
    h1 = gelu(z1)
    z2 = h1 @ p["W2"].T + p["b2"]
    h2 = gelu(z2)
    out = h2 @ p["W3"].T + p["b3"]
    return out, (x, z1, h1, z2, h2)


def loss_and_grads(p, x, t):
    """MSE over all B x D outputs and its gradient w.r.t. every parameter."""
    out, (x, z1, h1, z2, h2) = forward(p, x)
    diff = out - t
    loss = np.mean(diff**2)
    dout = 2.0 * diff / diff.size
    g = {"W3": dout.T @ h2, "b3": dout.sum(0)}
    dh2 = dout @ p["W3"]
    dz2 = dh2 * gelu_grad(z2)
    g["W2"] = dz2.T @ h1
    g["b2"] = dz2.sum(0)
    dh1 = dz2 @ p["W2"]
    dz1 = dh1 * gelu_grad(z1)
    g["W1"] = dz1.T @ x
    g["b1"] = dz1.sum(0)
    return loss, g


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015) in place; t is the 1-based step."""
    for k in p:
        m[k] = b1 * m[k] + (1 - b1) * g[k]
        v[k] = b2 * v[k] + (1 - b2) * g[k] ** 2
        mh = m[k] / (1 - b1**t)
        vh = v[k] / (1 - b2**t)
        p[k] -= lr * mh / (np.sqrt(vh) + eps)


def init_params(d, h, rng):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases."""
    def u(o, i):
        return rng.uniform(-1, 1, (o, i)) / np.sqrt(i)
    return {"W1": u(h, d), "b1": np.zeros(h), "W2": u(h, h), "b2": np.zeros(h),
            "W3": u(d, h), "b3": np.zeros(d)}
