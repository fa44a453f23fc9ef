"""Initialisation strategies and per-distribution initial values
(pyaugur/pyaugur/augurdistlib.py:23-169 of the reference), drawing from a
seeded numpy Generator instead of the global numpy state."""
import numpy as np
import scipy.stats as sps


class AugurInitStrat(object):
    pass


class AugurRandom(AugurInitStrat):
    """Draw initial values from the priors."""


class AugurDefault(AugurInitStrat):
    """Fixed default values (means / uniform / zeros)."""


class AugurUser(AugurInitStrat):
    """User-supplied point: dict name -> value."""

    def __init__(self, initPt):
        self.initPt = initPt


def initDirichlet(strat, alpha, rng):
    alpha = np.asarray(alpha, dtype=np.float64)
    if isinstance(strat, AugurDefault):
        return np.full(len(alpha), 1.0 / len(alpha))
    return rng.dirichlet(alpha)


def initCategorical(strat, pmf, rng, size=None):
    pmf = np.asarray(pmf, dtype=np.float64)
    if isinstance(strat, AugurDefault):
        return np.zeros(size, np.int32) if size is not None else 0
    cdf = np.cumsum(pmf)
    u = rng.uniform(0, cdf[-1], size=size)
    return np.minimum(np.searchsorted(cdf, u, side="right"), len(pmf) - 1).astype(np.int32)


def initNormal(strat, mean, variance, rng, size=None):
    if isinstance(strat, AugurDefault):
        return np.full(size, float(mean)) if size is not None else float(mean)
    return rng.normal(mean, np.sqrt(variance), size=size)


def initInverseGamma(strat, shape, scale, rng):
    if isinstance(strat, AugurDefault):
        return 0.5
    return 1.0 / rng.gamma(shape, 1.0 / scale)


def initExponential(strat, rate, rng):
    if isinstance(strat, AugurDefault):
        return 1.0
    return rng.exponential(1.0 / rate)


def initMvNormal(strat, mu, cov, rng):
    mu = np.asarray(mu, dtype=np.float64)
    if isinstance(strat, AugurDefault):
        return np.zeros(len(mu))
    return rng.multivariate_normal(mu, np.asarray(cov, dtype=np.float64))


def initInvWishart(strat, df, scale, rng):
    scale = np.asarray(scale, dtype=np.float64)
    if isinstance(strat, AugurDefault):
        return np.eye(scale.shape[0])
    return np.atleast_2d(sps.invwishart.rvs(df=df, scale=scale, random_state=rng))


def initGamma(strat, shape, rate, rng, size=None):
    if isinstance(strat, AugurDefault):
        return np.full(size, 0.5) if size is not None else 0.5
    return rng.gamma(shape, 1.0 / rate, size=size)


def initBeta(strat, a, b, rng, size=None):
    if isinstance(strat, AugurDefault):
        return np.full(size, 1.0) if size is not None else 1.0
    return rng.beta(a, b, size=size)


def initUniform(strat, lower, upper, rng, size=None):
    if isinstance(strat, AugurDefault):
        mid = (np.asarray(lower) + np.asarray(upper)) / 2.0
        return np.broadcast_to(mid, size).copy() if size is not None else float(mid)
    return rng.uniform(lower, upper, size=size)
