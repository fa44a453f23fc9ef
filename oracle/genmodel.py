"""Oracle for generic continuous models (test infrastructure).

Only tests, the smoke check and the benchmark's CPU baseline use this
module.  It checks results and is never the thing measured or shipped.

It restates on the CPU, with numpy, what the reference's compiler generates
for a model whose parameters are all continuous:

  log joint: every declaration's terms summed, in declaration order
    (compiler/augur/src/Core/CgLike.hs:90-267).  The per-distribution
    densities follow cbits/augur_dist.c:
      bernoulli :59-66, poisson :124-130, beta :158-165,
      exponential :231-238, gamma :265-273, invgamma :345-347,
      normal :372-381, uniform :426-433.
    Beta and Gamma are the logarithm of the reference's pdf taken term by
    term, which is the same function without its overflow for large
    shapes.  The Poisson's sum of log k for k = 1..pt is lgamma(pt + 1).
  gradient: the chain rule through +, -, *, /, sigmoid and dotprod, with
    the reference's per-argument derivatives (Core/CgGrad.hs:151-376):
      bernoulli_dotarg1 :68-71, poisson_dotarg1 :132-139,
      beta_dot* :167-177, exponential_dot* :240-256,
      gamma_dot* :275-285, invgamma_dot* :349-359,
      normal_dot* :383-394, uniform_dot* :435-455.
    Each term contributes dotpt * d(point)/dq + sum_i dotarg_i * d(arg_i)/dq.
    The derivatives of the point and argument expressions are taken by
    complex steps, so no hand-written adjoint is involved.  Those
    expressions are analytic, and complex steps give their derivatives to
    machine precision.
  DiscGibbs on finite discrete parameters (Categorical over a fixed pmf,
    Bernoulli): Core/CgDisc.hs:120-165, in log space (disc_step).
  HMC: cbits/augur_mcmclib.h:343-424, each block on its full conditional
    (the declarations that mention its variables, Core/RwCore.hs:58-67).
    The momentum and leapfrog updates are the reference's.  The step size
    and step count are jittered.  The Metropolis test is in FP64.  Random numbers come from the same Philox
    streams as the device: per (seed, chain), with counter
    (site, sub, sweep, tag).  Block b uses tags 16 + 4b + {0: momentum,
    1: jitter, 2: accept}.

Parameter layout: the parameters in declaration order, each flattened
row-major, give one vector q of P reals per chain.

Pinning:
  - the tests check ll against scipy.stats log densities;
  - they check the gradient against central differences of ll (both in the
    textbook mode);
  - they check the reference-compatible Normal forms against their closed
    forms;
  - they check the HMC restatement against a conjugate posterior.
"""
import math

import numpy as np
from scipy import special

from . import oracle as orc

LARGE_POS_NUM = 1e8  # cbits/augur_math.h:24
H = 1e-30           # complex step


# ---- densities and derivatives (cbits/augur_dist.c) -------------------------

def _where(cond, a, b):
    return np.where(cond, a, b)


def dist_ll(name, pt, args, ref_compat):
    a = args
    if name == "Categorical":  # cbits/augur_dist.c:100-104, pmf a fixed vector
        with np.errstate(divide="ignore"):
            return np.log(np.asarray(a[0], np.float64)[np.asarray(pt, np.int64)])
    if name == "Normal":
        mean, var = a
        t = pt - mean
        lv = (1.0 if ref_compat else 0.5) * np.log(np.where(var > 0, var, 1.0))
        return _where(var > 0, -lv - 0.5 * t * t / np.where(var > 0, var, 1.0), -LARGE_POS_NUM)
    if name == "Bernoulli":
        bias = a[0]
        ok = (bias >= 0.0) & (bias <= 1.0)
        return _where(ok, np.log(np.where(ok, pt * bias + (1 - pt) * (1.0 - bias), 1.0)), -LARGE_POS_NUM)
    if name == "Poisson":
        rate = a[0]
        pt_i = np.asarray(pt, np.int64)
        lf = special.gammaln(pt_i + 1.0)  # sum_{k=1}^{pt} log k
        return pt_i * np.log(rate) - rate - lf
    if name == "Exponential":
        rate = a[0]
        ok = (rate > 0.0) & (pt > 0.0)
        return _where(ok, np.log(np.where(ok, rate, 1.0)) - rate * pt, -LARGE_POS_NUM)
    if name == "Gamma":
        shape, rate = a
        return shape * np.log(rate) - special.gammaln(shape) + (shape - 1.0) * np.log(pt) - rate * pt
    if name == "InvGamma":
        shape, scale = a
        return shape * np.log(scale) - special.gammaln(shape) - (shape + 1.0) * np.log(pt) - scale / pt
    if name == "Beta":
        al, be = a
        return (special.gammaln(al + be) - special.gammaln(al) - special.gammaln(be)
                + (al - 1.0) * np.log(pt) + (be - 1.0) * np.log(1.0 - pt))
    if name == "Uniform":
        lo, hi = a
        ok = (lo < pt) & (pt < hi)
        return _where(ok, -np.log(np.where(ok, hi - lo, 1.0)), -LARGE_POS_NUM)
    raise ValueError(f"no density for {name}")


def dist_dot(name, which, pt, args, ref_compat):
    """d ll / d point (which = 'pt') or d ll / d argument i (which = i)."""
    a = args
    if name == "Normal":
        mean, var = a
        t = pt - mean
        if which == "pt":
            return -t / var
        if which == 0:
            return t / var
        return (-0.5 * var if ref_compat else -0.5 / var) + 0.5 * (t * t) / (var * var)
    if name == "Bernoulli":
        bias = a[0]
        return (pt + pt - 1) / (pt * bias + (1 - pt) * (1.0 - bias))
    if name == "Poisson":
        rate = a[0]
        return np.where(rate > 0.0, pt / np.where(rate > 0, rate, 1.0) - 1.0, -LARGE_POS_NUM)
    if name == "Exponential":
        rate = a[0]
        ok = (rate > 0.0) & (pt > 0.0)
        if which == "pt":
            return np.where(ok, -rate, 0.0)
        return np.where(ok, 1.0 / np.where(ok, rate, 1.0) - pt, 0.0)
    if name == "Gamma":
        shape, rate = a
        if which == "pt":
            return (shape - 1.0) / pt - rate
        if which == 0:
            return -special.digamma(shape) + np.log(rate) + np.log(pt)
        return shape / rate - pt
    if name == "InvGamma":
        shape, scale = a
        if which == "pt":
            return -(shape + 1.0) / pt + scale / (pt * pt)
        if which == 0:
            return np.log(scale) - special.digamma(shape) - np.log(pt)
        return shape / scale - 1.0 / pt
    if name == "Beta":
        al, be = a
        if which == "pt":
            return (al - 1.0) / pt - (be - 1.0) / (1.0 - pt)
        if which == 0:
            return special.digamma(al + be) - special.digamma(al) + np.log(pt)
        return special.digamma(al + be) - special.digamma(be) + np.log(1.0 - pt)
    if name == "Uniform":
        lo, hi = a
        ok = (lo < pt) & (pt < hi)
        if which == "pt":
            return np.zeros_like(np.asarray(pt, np.float64) + lo + hi)
        w = np.where(ok, 1.0 / np.where(ok, hi - lo, 1.0), 0.0)
        return w if which == 0 else -w
    raise ValueError(f"no derivative for {name}")


INT_DISTS = {"Bernoulli", "Poisson", "Categorical"}
DISC_DISTS = {"Bernoulli", "Categorical"}


# ---- model interpretation -------------------------------------------------------

def _free_vars(e, out):
    k = e[0]
    if k == "var":
        out.add(e[1])
    elif k == "proj":
        _free_vars(e[1], out)
        for i in e[2]:
            _free_vars(i, out)
    elif k in ("prim", "dist"):
        for a in e[2]:
            _free_vars(a, out)
    elif k == "bin":
        _free_vars(e[2], out)
        _free_vars(e[3], out)
    return out


class GenModel:
    """A parsed continuous model bound to its hyper-parameters and data.

    ``hypers`` and ``data`` map names to values (Python scalars or numpy
    arrays; a Vec (Vec Real) is given as a 2-D array)."""

    def __init__(self, model, hypers, data, ref_compat=True):
        self.model = model
        self.ref_compat = ref_compat
        self.env = {}
        for name, _ in model.hypers:
            v = hypers[name]
            self.env[name] = np.asarray(v) if not np.isscalar(v) else v
        for d in model.data:
            self.env[d.name] = np.asarray(data[d.name])
        self.layout = {}       # continuous: name -> (offset in q, shape)
        self.zlayout = {}      # discrete: name -> (offset in zq, shape)
        off = zoff = 0
        for d in model.params:
            shape = tuple(int(self._eval_const(hi)) for (_, lo, hi) in d.grids)
            size = int(np.prod(shape)) if shape else 1
            if d.exp[1] in DISC_DISTS:
                self.zlayout[d.name] = (zoff, shape)
                zoff += size
            else:
                self.layout[d.name] = (off, shape)
                off += size
        self.P = max(off, 1)
        self.NZ = max(zoff, 1)
        self.pnames = set(self.layout)
        self.zq = np.zeros(self.NZ, np.int64)  # discrete values used by ll / grad / hmc_step
        self.terms = [d for d in model.decls if d.stochastic]

    def _eval_const(self, e):
        v = self._ev(e, self.env)
        return v

    def unpack(self, q, zq=None):
        env = dict(self.env)
        for name, (off, shape) in self.layout.items():
            size = int(np.prod(shape)) if shape else 1
            v = q[off:off + size]
            env[name] = v.reshape(shape) if shape else v[0]
        zq = self.zq if zq is None else zq
        for name, (off, shape) in self.zlayout.items():
            size = int(np.prod(shape)) if shape else 1
            v = np.asarray(zq[off:off + size], np.int64)
            env[name] = v.reshape(shape) if shape else int(v[0])
        return env

    # expressions, vectorised over the grid variables bound in env
    def _ev(self, e, env):
        k = e[0]
        if k == "lit":
            return float(e[1]) if isinstance(e[1], float) else e[1]
        if k == "var":
            return env[e[1]]
        if k == "proj":
            base = self._ev(e[1], env)
            idx = [np.asarray(self._ev(i, env)).astype(np.int64) for i in e[2]]
            return base[tuple(idx)] if len(idx) > 1 else base[idx[0]]
        if k == "bin":
            a, b = self._ev(e[2], env), self._ev(e[3], env)
            return {"+": lambda: a + b, "-": lambda: a - b, "*": lambda: a * b, "/": lambda: a / b}[e[1]]()
        if k == "prim":
            if e[1] == "sigmoid":
                t = self._ev(e[2][0], env)
                return 1.0 / (1.0 + np.exp(-t))
            if e[1] == "dotprod":
                a, b = self._ev(e[2][0], env), self._ev(e[2][1], env)
                return np.sum(a * b, axis=-1)
        raise ValueError(f"cannot evaluate {e}")

    def _grid_env(self, env, d):
        if not d.grids:
            return env, 1
        sizes = [int(self._ev(hi, env)) for (_, lo, hi) in d.grids]
        idx = np.meshgrid(*[np.arange(s) for s in sizes], indexing="ij")
        g = dict(env)
        for (var, _, _), ix in zip(d.grids, idx):
            g[var] = ix.ravel()
        return g, int(np.prod(sizes))

    def _term_parts(self, d, env):
        g, n = self._grid_env(env, d)
        pt_e = ("proj", ("var", d.name), [("var", i) for i in d.idxs]) if d.idxs else ("var", d.name)
        pt = self._ev(pt_e, g)
        args = [self._ev(a, g) for a in d.exp[2]]
        return pt_e, pt, args, n

    def ll_terms(self, q):
        """Per-declaration log-density sums at the point q (FP64)."""
        env = self.unpack(np.asarray(q, np.float64))
        out = []
        for d in self.terms:
            _, pt, args, n = self._term_parts(d, env)
            out.append(float(np.sum(np.broadcast_to(dist_ll(d.exp[1], pt, args, self.ref_compat), (n,)))))
        return out

    def ll(self, q, block_vars=None):
        """Log joint, or with block_vars the block's full conditional: the
        terms of the declarations that mention a block variable, each once
        (compiler/augur/src/Core/RwCore.hs:58-67)."""
        terms = self.ll_terms(q)
        if block_vars is None:
            return float(sum(terms))
        vs = set(block_vars)
        return float(sum(t for d, t in zip(self.terms, terms)
                         if d.name in vs or (_free_vars(d.exp, set()) & vs)))

    def grad(self, q):
        """The reference's gradient of the log joint with respect to q."""
        q = np.asarray(q, np.float64)
        env = self.unpack(q)
        g = np.zeros(self.P)
        for d in self.terms:
            pt_e, pt, args, n = self._term_parts(d, env)
            pieces = [("pt", pt_e)] + list(enumerate(d.exp[2]))
            dots = {}
            for which, e in pieces:
                if not (_free_vars(e, set()) & self.pnames):
                    continue
                dots[which] = np.broadcast_to(dist_dot(d.exp[1], which, pt, args, self.ref_compat), (n,))
            if not dots:
                continue
            for j in range(self.P):
                qc = q.astype(np.complex128)
                qc[j] += 1j * H
                envc = self.unpack(qc)
                gc, _ = self._grid_env(envc, d)
                s = 0.0
                for which, e in pieces:
                    if which not in dots:
                        continue
                    de = np.broadcast_to(np.imag(np.asarray(self._ev(e, gc), np.complex128)) / H, (n,))
                    s += float(np.sum(dots[which] * de))
                g[j] += s
        return g

    # ---- HMC (cbits/augur_mcmclib.h:343-424) ----
    def hmc_step(self, q, coords, sim_len, eps, fixed_steps, seed, chain, sweep, block, block_vars=None):
        """One HMC update of the coordinates `coords` (a boolean mask) of q,
        accepted on the block's full conditional when block_vars names the
        block's variables (else on the log joint); returns (new q,
        accepted)."""
        tag = 16 + 4 * block
        step, L = jitter(seed, chain, sweep, tag + 1, sim_len, eps, fixed_steps)
        mask = np.asarray(coords, bool)
        p0 = np.zeros(self.P)
        for j in np.nonzero(mask)[0]:
            w = orc.draw4(seed, chain, int(j), 0, sweep, tag)
            p0[j] = normal_from(int(w[0]), int(w[1]))
        g = self.grad(q)
        p = p0 + 0.5 * step * g * mask
        x = q.copy()
        for l in range(L):
            x = x + step * p * mask
            g = self.grad(x)
            coef = step if l != L - 1 else 0.5 * step
            p = p + coef * g * mask
        k0 = 0.5 * float(np.sum(p0[mask] ** 2))
        k1 = 0.5 * float(np.sum(p[mask] ** 2))
        w = orc.draw4(seed, chain, 0, 0, sweep, tag + 2)
        u = u01(int(w[0]))
        acc = math.log(u) < (self.ll(x, block_vars) - self.ll(q, block_vars)) + (k0 - k1)
        return (x if acc else q.copy()), acc


def disc_step(model, q, zq, name, seed, chain, sweep, block):
    """DiscGibbs on the discrete parameter `name` (Core/CgDisc.hs:120-165):
    at each point of its grid the log full conditional of every value (its
    prior and the terms that use it at that point, all points at once), then
    the inverse-CDF draw in log space with the maximum subtracted, from the
    uniform of Philox counter (point, 0, sweep, 16 + 4 block) under key
    (seed, chain).  Returns the new discrete vector."""
    off, shape = model.zlayout[name]
    n = int(np.prod(shape)) if shape else 1
    d0 = model.model.decl(name)
    K = 2 if d0.exp[1] == "Bernoulli" else len(np.asarray(model.env[d0.exp[2][0][1]]))
    terms = [d for d in model.terms if d.name == name or (name in _free_vars(d.exp, set()))]
    q = np.asarray(q, np.float64)
    lp = np.zeros((n, K))
    for k in range(K):
        zk = np.array(zq, np.int64)
        zk[off:off + n] = k
        env = model.unpack(q, zk)
        for d in terms:
            _, pt, args, m = model._term_parts(d, env)
            lp[:, k] += np.broadcast_to(dist_ll(d.exp[1], pt, args, model.ref_compat), (m,))
    out = np.array(zq, np.int64)
    tag = 16 + 4 * block
    for f in range(n):
        w = np.exp(lp[f] - lp[f].max())
        tot = 0.0
        for v in w:
            tot += v
        u = u01(int(orc.draw4(seed, chain, f, 0, sweep, tag)[0])) * tot
        pick, cum = K - 1, 0.0
        for k in range(K):
            cum += w[k]
            if u < cum:
                pick = k
                break
        out[off + f] = pick
    return out


def u01(w):
    return (w + 0.5) * 2.3283064365386963e-10


def normal_from(w0, w1):
    u1, u2 = u01(w0), u01(w1)
    return math.sqrt(-2.0 * math.log(u1)) * math.cos(6.283185307179586476925 * u2)


def jitter(seed, chain, sweep, tag, sim_len, eps, fixed_steps):
    """Step size eps U(0.8, 1.2) and step count max(1, round(U(0.9, 1.1)
    simLen / step)) (cbits/augur_mcmclib.h:401-405); fixed_steps > 0 turns
    the jitter off, fixed_steps < 0 keeps the step jitter with -fixed_steps
    steps."""
    if fixed_steps > 0:
        return eps, fixed_steps
    w = orc.draw4(seed, chain, 0, 0, sweep, tag)
    st = (0.8 + 0.4 * u01(int(w[0]))) * eps
    if fixed_steps < 0:
        return st, -fixed_steps
    sc = 0.9 + 0.2 * u01(int(w[1]))
    r = math.floor(sc * sim_len / st + 0.5)  # C round(): half away from zero
    return st, max(1, int(r))
