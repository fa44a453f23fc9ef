"""Models shared by the generic-lowering tests (CPU pins and GPU parity):
continuous models outside the three families, with data generators and
in-support points."""
import numpy as np

LINREG = """(K : Int, N : Int, a : Real, b : Real, x : Vec (Vec Real)) => {
  param s2 ~ InvGamma(a, b) ;
  param w[k] ~ Normal(0.0, s2)
      for k <- 0 until K ;
  data y[n] ~ Normal(dotprod(w, x[n]), s2)
      for n <- 0 until N ;
}
"""

HLR = """(K : Int, N : Int, lam : Real, x : Vec (Vec Real)) => {
  param v ~ Exponential(lam) ;
  param b ~ Normal(0.0, v) ;
  param theta[k] ~ Normal(0.0, v)
      for k <- 0 until K ;
  data y[n] ~ Bernoulli( sigmoid(dotprod(theta, x[n]) + b) )
      for n <- 0 until N ;
}
"""

POIS = """(G : Int, N : Int, a : Real, r : Real, grp : Vec Int) => {
  param shape ~ Uniform(0.5, 4.0) ;
  param lam[g] ~ Gamma(shape, r)
      for g <- 0 until G ;
  data y[n] ~ Poisson(lam[grp[n]])
      for n <- 0 until N ;
}
"""

BETA = """(N : Int, a : Real, b : Real) => {
  param p ~ Beta(a, b) ;
  param t ~ Exponential(2.0) ;
  data y[n] ~ Bernoulli(p)
      for n <- 0 until N ;
  data u[n] ~ Exponential(t * 2.0)
      for n <- 0 until N ;
}
"""

MIXED = """(G : Int, K : Int, N : Int, grp : Vec Int, x : Vec (Vec Real)) => {
  param tau ~ InvGamma(3.0, 2.0) ;
  param w[g, k] ~ Normal(0.0, tau)
      for g <- 0 until G, k <- 0 until K ;
  data y[n] ~ Normal(dotprod(w[grp[n]], x[n]) / (1.0 + tau), tau * 0.5)
      for n <- 0 until N ;
}
"""


def models(rng, n=60):
    K, G = 4, 3
    x = rng.standard_normal((n, K))
    grp = rng.integers(0, G, n).astype(np.int32)
    return {
        "linreg": (LINREG, dict(K=K, N=n, a=3.0, b=2.0, x=x), dict(y=x @ rng.standard_normal(K)
                                                                 + rng.standard_normal(n))),
        "hlr": (HLR, dict(K=K, N=n, lam=1.0, x=x), dict(y=(rng.random(n) < 0.5).astype(np.int32))),
        "pois": (POIS, dict(G=G, N=n, a=2.0, r=1.5, grp=grp), dict(y=rng.poisson(2.0, n).astype(np.int32))),
        "beta": (BETA, dict(N=n, a=2.0, b=3.0), dict(y=(rng.random(n) < 0.3).astype(np.int32),
                                                     u=rng.exponential(0.5, n))),
        "mixed": (MIXED, dict(G=G, K=K, N=n, grp=grp, x=x), dict(y=rng.standard_normal(n))),
    }


def point(m, rng):
    """An in-support point for the model's parameters."""
    q = np.zeros(m.P)
    for name, (off, shape) in m.layout.items():
        size = int(np.prod(shape)) if shape else 1
        d = m.model.decl(name).exp[1]
        if d in ("Exponential", "Gamma", "InvGamma"):
            q[off:off + size] = rng.uniform(0.5, 2.0, size)
        elif d == "Beta":
            q[off:off + size] = rng.uniform(0.2, 0.8, size)
        elif d == "Uniform":
            q[off:off + size] = rng.uniform(1.0, 3.0, size)
        else:
            q[off:off + size] = rng.normal(0.0, 0.7, size)
    return q


MIX = """(K : Int, N : Int, pis : Vec Real, s2 : Real) => {
  param mu[k] ~ Normal(0.0, 10.0)
      for k <- 0 until K ;
  param z[n] ~ Categorical(pis)
      for n <- 0 until N ;
  data y[n] ~ Normal(mu[z[n]], s2)
      for n <- 0 until N ;
}
"""

SPIKE = """(N : Int) => {
  param p ~ Beta(1.0, 1.0) ;
  param z[n] ~ Bernoulli(p)
      for n <- 0 until N ;
  data y[n] ~ Normal(z[n] * 3.0, 1.0)
      for n <- 0 until N ;
}
"""


def mixture(rng, n=600, centers=(-4.0, 0.0, 5.0), s2=1.0):
    """A 1-D mixture with fixed weights and variance: hypers, data and the
    generating assignment."""
    K = len(centers)
    pis = np.full(K, 1.0 / K)
    z = rng.integers(0, K, n)
    y = np.asarray(centers)[z] + rng.normal(0.0, np.sqrt(s2), n)
    return dict(K=K, N=n, pis=pis, s2=s2), dict(y=y), z.astype(np.int32)
