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
    # Geyer initial positive sequence on pairs (rho[2t] + rho[2t+1] up to the
    # first negative pair), then monotone
    npairs = n // 2
    pairs = rho[0:2 * npairs:2] + rho[1:2 * npairs:2]
    neg = np.flatnonzero(pairs < 0)
    pair_sums = pairs[:neg[0]] if neg.size else pairs
    if not pair_sums.size:
        return float(m * n)
    p = np.minimum.accumulate(pair_sums)
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


def _split_cols(d):
    """(iterations, chains, k) -> (2 chains, iterations // 2, k): each chain's
    first and last halves as separate chains (as _split, per coordinate)."""
    x = np.moveaxis(np.asarray(d, dtype=np.float64), 1, 0)
    n = x.shape[1] // 2
    return np.concatenate([x[:, :n], x[:, x.shape[1] - n:]], axis=0)


def _rank_normalize_cols(x):
    """_rank_normalize of each coordinate x[:, :, j] (average ranks of ties,
    so the order a sort leaves ties in does not matter)."""
    m, n, k = x.shape
    total = m * n
    flat = np.ascontiguousarray(x.reshape(total, k).T)  # (k, total)
    order = flat.argsort(axis=1)
    sv = np.take_along_axis(flat, order, axis=1)
    # runs of equal values in each sorted row: ranks start + 1 .. end + 1
    newrun = np.ones((k, total), dtype=bool)
    newrun[:, 1:] = sv[:, 1:] != sv[:, :-1]
    pos = np.arange(total)[None, :]
    start = np.maximum.accumulate(np.where(newrun, pos, 0), axis=1)
    endrun = np.ones((k, total), dtype=bool)
    endrun[:, :-1] = newrun[:, 1:]
    end = np.minimum.accumulate(np.where(endrun, pos, total)[:, ::-1], axis=1)[:, ::-1]
    ranks = np.empty((k, total))
    np.put_along_axis(ranks, order, (start + end + 2) / 2.0, axis=1)
    z = ndtri((ranks - 0.375) / (total + 0.25))
    return np.ascontiguousarray(z.T).reshape(m, n, k)


def _rhat_raw_cols(x):
    m, n, k = x.shape
    if n < 2:
        return np.full(k, np.nan)
    w = x.var(axis=1, ddof=1).mean(axis=0)
    b = n * x.mean(axis=1).var(axis=0, ddof=1) if m > 1 else np.zeros(k)
    var_plus = (n - 1) / n * w + b / n
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.sqrt(var_plus / w)
    return np.where(w == 0, np.where(b == 0, np.nan, np.inf), r)


def _ess_raw_cols(x):
    """_ess_raw of each coordinate x[:, :, j]."""
    m, n, k = x.shape
    if n < 4:
        return np.full(k, np.nan)
    xc = np.ascontiguousarray(np.moveaxis(x - x.mean(axis=1, keepdims=True), 1, 2))  # (m, k, n)
    size = 1
    while size < 2 * n:
        size *= 2
    f = np.fft.rfft(xc, n=size, axis=2)
    acov = np.fft.irfft(f * np.conj(f), n=size, axis=2)[:, :, :n] / n  # (m, k, n)
    mean_var = acov[:, :, 0].mean(axis=0) * n / (n - 1.0)
    var_plus = mean_var * (n - 1.0) / n
    if m > 1:
        var_plus = var_plus + x.mean(axis=1).var(axis=0, ddof=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        rho = (1.0 - (mean_var[:, None] - acov.mean(axis=0)) / var_plus[:, None]).T  # (n, k)
    rho[0] = 1.0
    npairs = n // 2
    pairs = rho[0:2 * npairs:2] + rho[1:2 * npairs:2]  # (npairs, k)
    neg = pairs < 0
    first = np.where(neg.any(axis=0), neg.argmax(axis=0), npairs)
    p = np.minimum.accumulate(np.where(np.arange(npairs)[:, None] < first, pairs, np.inf), axis=0)
    psum = np.where(np.arange(npairs)[:, None] < first, p, 0.0).sum(axis=0)
    tau = np.maximum(-1.0 + 2.0 * psum, 1.0 / np.log10(m * n))
    with np.errstate(divide="ignore", invalid="ignore"):
        ess = m * n / tau
    ess = np.where(first == 0, float(m * n), ess)
    return np.where(var_plus <= 0, np.nan, ess)


def summarize(draws):
    """draws: (iterations, chains, k) -> dict of per-coordinate arrays: mean,
    rhat (max of bulk and tail), ess_bulk, ess_tail, mcse (of the mean).
    The same quantities as rhat / ess_bulk / ess_tail / mcse_mean on each
    coordinate, computed for all coordinates at once."""
    d = np.asarray(draws, dtype=np.float64)
    s = _split_cols(d)
    z = _rank_normalize_cols(s)
    med = np.median(s.reshape(-1, s.shape[2]), axis=0)
    bulk = _rhat_raw_cols(z)
    tail = _rhat_raw_cols(_rank_normalize_cols(np.abs(s - med)))
    flat = s.reshape(-1, s.shape[2])
    q05, q95 = np.quantile(flat, [0.05, 0.95], axis=0)
    e1 = _ess_raw_cols((s <= q05).astype(np.float64))
    e2 = _ess_raw_cols((s <= q95).astype(np.float64))
    e_mean = _ess_raw_cols(s)
    allx = np.moveaxis(d, 1, 0).reshape(-1, d.shape[2])
    with np.errstate(divide="ignore", invalid="ignore"):
        mcse = np.where(np.isfinite(e_mean) & (e_mean != 0), allx.std(axis=0, ddof=1) / np.sqrt(e_mean), np.nan)
    return {"mean": d.mean(axis=(0, 1)), "rhat": np.maximum(bulk, tail), "ess_bulk": _ess_raw_cols(z),
            "ess_tail": np.minimum(e1, e2), "mcse": mcse}
