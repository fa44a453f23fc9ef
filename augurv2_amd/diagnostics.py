"""MCMC convergence and efficiency diagnostics.

Rank-normalised split-R-hat and bulk / tail effective sample size of
Vehtari, Gelman, Simpson, Carpenter, Buerkner, "Rank-normalization,
folding, and localization: an improved R-hat for assessing convergence of
MCMC", Bayesian Analysis 16(2), 2021, with Geyer's initial monotone
sequence estimator of the integrated autocorrelation time.  The reference
computes none of these (it only times sampling), so they are the
measurement layer of this runtime.

All functions take draws shaped (chains, draws) for one scalar quantity.
"""
import numpy as np
from scipy.special import ndtri


def _split(x):
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 1:
        x = x[None, :]
    n = x.shape[1] // 2
    return np.concatenate([x[:, :n], x[:, x.shape[1] - n:]], axis=0)


def _rank_normalize(x):
    flat = x.ravel()
    order = flat.argsort(kind="mergesort")
    ranks = np.empty(flat.size)
    # average ranks for ties
    sorted_vals = flat[order]
    r = np.arange(1, flat.size + 1, dtype=np.float64)
    uniq, start, counts = np.unique(sorted_vals, return_index=True, return_counts=True)
    avg = (2 * start + counts + 1) / 2.0
    r = np.repeat(avg, counts)
    ranks[order] = r
    z = ndtri((ranks - 0.375) / (flat.size + 0.25))
    return z.reshape(x.shape)


def _rhat_raw(x):
    m, n = x.shape
    if n < 2:
        return np.nan
    chain_mean = x.mean(axis=1)
    chain_var = x.var(axis=1, ddof=1)
    w = chain_var.mean()
    b = n * chain_mean.var(ddof=1) if m > 1 else 0.0
    if w == 0:
        return np.nan if b == 0 else np.inf
    var_plus = (n - 1) / n * w + b / n
    return float(np.sqrt(var_plus / w))


def _autocov(x):
    """Biased autocovariance of each row (FFT)."""
    m, n = x.shape
    xc = x - x.mean(axis=1, keepdims=True)
    size = 1
    while size < 2 * n:
        size *= 2
    f = np.fft.rfft(xc, n=size, axis=1)
    ac = np.fft.irfft(f * np.conj(f), n=size, axis=1)[:, :n]
    return ac / n


def _ess_raw(x):
    m, n = x.shape
    if n < 4:
        return np.nan
    acov = _autocov(x)
    chain_mean = x.mean(axis=1)
    mean_var = acov[:, 0].mean() * n / (n - 1.0)
    var_plus = mean_var * (n - 1.0) / n
    if m > 1:
        var_plus += chain_mean.var(ddof=1)
    if var_plus <= 0:
        return np.nan
    rho = 1.0 - (mean_var - acov.mean(axis=0)) / var_plus
    rho[0] = 1.0
    # Geyer initial positive sequence on pairs, then monotone
    t = 0
    pair_sums = []
    while t + 1 < n:
        p = rho[t] + rho[t + 1]
        if p < 0:
            break
        pair_sums.append(p)
        t += 2
    if not pair_sums:
        return float(m * n)
    p = np.minimum.accumulate(np.array(pair_sums))
    tau = -1.0 + 2.0 * p.sum()
    tau = max(tau, 1.0 / np.log10(m * n))
    return float(m * n / tau)


def rhat(x):
    """max(bulk, tail) rank-normalised split R-hat."""
    s = _split(x)
    bulk = _rhat_raw(_rank_normalize(s))
    folded = np.abs(s - np.median(s))
    tail = _rhat_raw(_rank_normalize(folded))
    return float(max(bulk, tail))


def ess_bulk(x):
    return _ess_raw(_rank_normalize(_split(x)))


def ess_tail(x):
    s = _split(x)
    q05, q95 = np.quantile(s, [0.05, 0.95])
    e1 = _ess_raw((s <= q05).astype(np.float64))
    e2 = _ess_raw((s <= q95).astype(np.float64))
    return float(min(e1, e2))


def ess_mean(x):
    """ESS for the mean (no rank normalisation), used for MCSE."""
    return _ess_raw(_split(x))


def mcse_mean(x):
    x = np.asarray(x, dtype=np.float64)
    e = ess_mean(x)
    return float(x.std(ddof=1) / np.sqrt(e)) if e and np.isfinite(e) else np.nan


def summarize(draws):
    """draws: (iterations, chains, k) -> dict of per-coordinate arrays."""
    d = np.asarray(draws, dtype=np.float64)
    k = d.shape[2]
    out = {"mean": d.mean(axis=(0, 1)), "rhat": np.empty(k), "ess_bulk": np.empty(k),
           "ess_tail": np.empty(k), "mcse": np.empty(k)}
    for j in range(k):
        xj = d[:, :, j].T
        out["rhat"][j] = rhat(xj)
        out["ess_bulk"][j] = ess_bulk(xj)
        out["ess_tail"][j] = ess_tail(xj)
        out["mcse"][j] = mcse_mean(xj)
    return out
