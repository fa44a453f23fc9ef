"""Seeded synthetic workloads of the five benchmark configurations.

Each generator follows the model's generative process with numpy
``default_rng(seed)``; no dataset is downloaded.
"""
import numpy as np


def lda_corpus(K=10, V=1000, D=200, mean_len=100, alpha=0.1, beta=0.01, seed=0):
    """LDA generative model: theta_d ~ Dir(alpha), phi_k ~ Dir(beta),
    N_d ~ Poisson(mean_len) (>= 1), z ~ Cat(theta_d), w ~ Cat(phi_z).
    Returns (doc_off[D + 1], w[T], alpha[K], beta[V], z_true[T])."""
    rng = np.random.default_rng(seed)
    theta = rng.dirichlet(np.full(K, alpha), D)
    phi = rng.dirichlet(np.full(V, beta), K)
    lens = np.maximum(rng.poisson(mean_len, D), 1)
    doc_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    z = np.concatenate([rng.choice(K, n, p=theta[d]) for d, n in enumerate(lens)]).astype(np.int32)
    cdf = np.cumsum(phi, axis=1)
    u = rng.random(len(z))
    w = np.minimum((cdf[z] < u[:, None]).sum(axis=1), V - 1).astype(np.int32)
    return doc_off, w, np.full(K, alpha), np.full(V, beta), z


def hlr_data(N=100_000, D=64, seed=0, dtype=np.float32):
    """X ~ N(0, 1); v = tau^2 ~ InvGamma(1, 1); beta_j ~ N(0, v);
    y ~ Bernoulli(sigmoid(X beta)).  Returns (x, y, v, beta)."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((N, D), dtype=np.float32).astype(dtype)
    v = 1.0 / rng.gamma(1.0, 1.0)
    beta = rng.standard_normal(D) * np.sqrt(v)
    t = x.astype(np.float64) @ beta
    with np.errstate(over="ignore"):  # exp(-t) = inf gives probability 0, as intended
        y = (rng.random(N) < 1.0 / (1.0 + np.exp(-t))).astype(np.int32)
    return x, y, v, beta


def hlr_init(D, chains, seed=1):
    """AugurRandom-style start per chain: v ~ InvGamma(1, 1), theta ~ N(0, v)
    (pyaugur/pyaugur/augurdistlib.py:119-133), scaled down so that the first
    trajectories start inside the region the step size resolves."""
    out = []
    for c in range(chains):
        rng = np.random.default_rng(seed + c)
        v = 1.0 / rng.gamma(1.0, 1.0)
        theta = rng.standard_normal(D) * 0.1
        out.append(np.concatenate([theta, [0.0], [v]]))
    return out


def batched_hlr_data(n_datasets=256, n_min=1000, n_max=200_000, dims=(16, 32, 64), seed=0):
    """Datasets for the batched HLR sampler: N_i ~ U{n_min..n_max}, D_i
    uniform over dims, each generated as hlr_data with seed 1000 + i.
    Returns (xs, ys)."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(n_min, n_max + 1, n_datasets)
    ds = rng.choice(np.asarray(dims), n_datasets)
    xs, ys = [], []
    for i in range(n_datasets):
        x, y, _, _ = hlr_data(int(ns[i]), int(ds[i]), seed=1000 + i)
        xs.append(x)
        ys.append(y)
    return xs, ys


def gmm_data(N=1_000_000, D=16, K=32, seed=0):
    """pis ~ Dir(1), mu_k ~ N(0, 10 I), cov_k ~ InvWishart(D + 2, I),
    z ~ Cat(pis), y ~ N(mu_z, cov_z).  Returns (y f32, z, pis, mu, cov)."""
    rng = np.random.default_rng(seed)
    pis = rng.dirichlet(np.ones(K))
    mu = rng.standard_normal((K, D)) * np.sqrt(10.0)
    cov = np.empty((K, D, D))
    df = D + 2
    for k in range(K):
        # Bartlett: W ~ Wishart(df, I), cov = W^-1
        a = np.zeros((D, D))
        for i in range(D):
            a[i, i] = np.sqrt(rng.chisquare(df - i))
            a[i, :i] = rng.standard_normal(i)
        w = a @ a.T
        cov[k] = np.linalg.inv(w)
    z = rng.choice(K, N, p=pis)
    y = np.empty((N, D), np.float32)
    for k in range(K):
        idx = np.nonzero(z == k)[0]
        if len(idx):
            L = np.linalg.cholesky(cov[k])
            y[idx] = (mu[k] + rng.standard_normal((len(idx), D)) @ L.T).astype(np.float32)
    return y, z.astype(np.int32), pis, mu, cov
