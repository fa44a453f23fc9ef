"""Kernel-level entry points of the C-ABI (numpy in, numpy out).

Each call runs one HIP kernel of the sampler so it can be compared with
the CPU restatement in isolation.
"""
import ctypes as C

import numpy as np

from . import _lib as L


def philox(ctr, key):
    ctr = L.u32(ctr).reshape(-1, 4)
    key = L.u32(key).reshape(-1, 2)
    n = ctr.shape[0]
    out = np.zeros((n, 4), np.uint32)
    L.call("augur_op_philox", L.ptr(ctr), L.ptr(key), n, L.ptr(out))
    return out


def gamma(seed, chain, sweep, tag, shapes):
    s = L.f64(shapes)
    out = np.zeros_like(s)
    L.call("augur_op_gamma", seed, chain, sweep, tag, L.ptr(s), len(s), L.ptr(out))
    return out


def dirichlet_conj(seed, chain, sweep, tag, alpha, cnt):
    a = L.f64(alpha)
    c = L.i32(cnt).reshape(-1, len(a))
    out = np.zeros(c.shape, np.float64)
    L.call("augur_op_dirichlet_conj", seed, chain, sweep, tag, L.ptr(a), L.ptr(c), c.shape[0],
           len(a), L.ptr(out))
    return out


def lda_zstep(doc_off, w, theta32, phi_t32, n_vocab, seed, chain, sweep):
    doc_off, w = L.i32(doc_off), L.i32(w)
    th, ph = L.f32(theta32), L.f32(phi_t32)
    D, K = th.shape
    z = np.zeros(len(w), np.int32)
    ct = np.zeros((D, K), np.int32)
    cp = np.zeros((K, n_vocab), np.int32)
    L.call("augur_op_lda_zstep", D, K, n_vocab, L.ptr(doc_off), L.ptr(w), L.ptr(th), L.ptr(ph),
           seed, chain, sweep, L.ptr(z), L.ptr(ct), L.ptr(cp))
    return z, ct, cp


def hlr_ll_grad(x, y, beta, has_intercept=False):
    """beta: chains x (D + 1) (theta, b).  Returns (ll[chains], grad[chains, D + 1])."""
    x, y = L.f32(x), L.i32(y)
    n, d = x.shape
    beta = L.f64(beta).reshape(-1, d + 1)
    c = beta.shape[0]
    ll = np.zeros(c)
    g = np.zeros((c, d + 1))
    L.call("augur_op_hlr_ll_grad", n, d, c, L.ptr(x), L.ptr(y), L.ptr(beta), int(has_intercept),
           L.ptr(ll), L.ptr(g))
    return ll, g


def hlr_ll64(x, y, beta, has_intercept=False):
    """The data log-density in FP64 row arithmetic.  beta: chains x (D + 1)."""
    x, y = L.f32(x), L.i32(y)
    n, d = x.shape
    beta = L.f64(beta).reshape(-1, d + 1)
    ll = np.zeros(beta.shape[0])
    L.call("augur_op_hlr_ll64", n, d, beta.shape[0], L.ptr(x), L.ptr(y), L.ptr(beta),
           int(has_intercept), L.ptr(ll))
    return ll


def gmm_zstep(y, pis, mu, cov, seed, chain, sweep):
    y = L.f32(y)
    n, d = y.shape
    pis, mu, cov = L.f64(pis), L.f64(mu), L.f64(cov)
    z = np.zeros(n, np.int32)
    L.call("augur_op_gmm_zstep", n, d, len(pis), L.ptr(y), L.ptr(pis), L.ptr(mu), L.ptr(cov),
           seed, chain, sweep, L.ptr(z))
    return z


def gmm_conj(alpha, mu0, cov0, df, scale0, cnt, sum_y, sum_yy, pis, mu, cov, seed, sweep,
             sample_pi=True, sample_cov=True):
    alpha, mu0, cov0, scale0 = L.f64(alpha), L.f64(mu0), L.f64(cov0), L.f64(scale0)
    k, d = len(alpha), len(mu0)
    cnt = L.i32(cnt)
    sum_y, sum_yy = L.f64(sum_y), L.f64(sum_yy)
    pis, mu, cov = L.f64(pis).copy(), L.f64(mu).copy(), L.f64(cov).copy()
    L.call("augur_op_gmm_conj", d, k, L.ptr(alpha), L.ptr(mu0), L.ptr(cov0), int(df),
           L.ptr(scale0), L.ptr(cnt), L.ptr(sum_y), L.ptr(sum_yy), int(sample_pi),
           int(sample_cov), seed, 0, sweep, L.ptr(pis), L.ptr(mu), L.ptr(cov))
    return pis, mu, cov


def mlp_loss_grad(params, x, t, d, h):
    """Surrogate MLP: outputs, MSE loss and parameter gradient (FP32 kernels)."""
    params, x, t = L.f32(params), L.f32(x), L.f32(t)
    rows = x.shape[0]
    out = np.zeros((rows, d), np.float32)
    loss = np.zeros(1)
    grads = np.zeros_like(params)
    L.call("augur_op_mlp_loss_grad", d, h, rows, L.ptr(params), L.ptr(x), L.ptr(t), L.ptr(out),
           L.ptr(loss), L.ptr(grads))
    return out, float(loss[0]), grads


def mlp_train(params, x, t, d, h, steps, batch, lr, seed):
    """Adam over Philox-drawn minibatches; returns (params, loss per step)."""
    p = L.f32(params).copy()
    x, t = L.f32(x), L.f32(t)
    hist = np.zeros(steps)
    L.call("augur_op_mlp_train", d, h, x.shape[0], L.ptr(x), L.ptr(t), int(steps), int(batch),
           float(lr), seed & 0xFFFFFFFF, L.ptr(p), L.ptr(hist))
    return p, hist
