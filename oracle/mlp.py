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


def bf16_round(a):
    """Each value rounded to the nearest BF16 (8 significand bits), ties to
    even, as FP64: the rounding of IEEE 754 round-to-nearest-even applied to
    the FP32 value, done on its bit pattern (add 0x7FFF plus the lowest kept
    bit, clear the low 16 bits)."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


def _ident(a):
    return a


def forward(p, x, rnd=None):
    """rnd: applied to both operands of every product (None: exact), e.g.
    bf16_round for the device's matrix-core training products."""
    r = rnd or _ident
    z1 = r(x) @ r(p["W1"]).T + p["b1"]
    h1 = gelu(z1)
    z2 = r(h1) @ r(p["W2"]).T + p["b2"]
    h2 = gelu(z2)
    out = r(h2) @ r(p["W3"]).T + p["b3"]
    return out, (x, z1, h1, z2, h2)


def loss_and_grads(p, x, t, rnd=None):
    """MSE over all B x D outputs and its gradient w.r.t. every parameter;
    rnd as forward (the bias sums are not products and stay exact)."""
    r = rnd or _ident
    out, (x, z1, h1, z2, h2) = forward(p, x, rnd)
    diff = out - t
    loss = np.mean(diff**2)
    dout = 2.0 * diff / diff.size
    g = {"W3": r(dout).T @ r(h2), "b3": dout.sum(0)}
    dh2 = r(dout) @ r(p["W3"])
    dz2 = dh2 * gelu_grad(z2)
    g["W2"] = r(dz2).T @ r(h1)
    g["b2"] = dz2.sum(0)
    dh1 = r(dz2) @ r(p["W2"])
    dz1 = dh1 * gelu_grad(z1)
    g["W1"] = r(dz1).T @ r(x)
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
