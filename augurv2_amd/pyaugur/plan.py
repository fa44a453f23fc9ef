"""Structural matching of a parsed model + schedule onto the sampler
families this runtime executes on the GPU, and the conversions between the
reference's Python value containers and the C-ABI buffers.

Families (each recognised up to variable names):
  LDA  -- compiler/augur/test/lda.rv with ConjGibbs theta, phi, DiscGibbs z
  HLR  -- compiler/augur/test/hlr.rv family (InvGamma or Exponential prior on
          the variance, optional intercept, the Bernoulli bias optionally a
          deterministic parameter as in dethlr.rv) with HMC blocks and MWG on
          v or b
  GMM  -- compiler/augur/test/hmvgmm.rv (and mvgmm.rv: fixed pis and covs)
          with ConjGibbs pis, ConjGibbs | ESlice | HMC | MWG | RSlice | NUTS mu,
          ConjGibbs cov, DiscGibbs z
  GEN  -- any other model whose parameters are continuous or finite discrete
          (Categorical over a fixed pmf, Bernoulli): the passes are generated
          from the model (pyaugur/lower.py, compiled at run time with hipRTC)
          and run as HMC and DiscGibbs blocks.
The reference's compiler generates code for any model (Compile.hs:311-345);
models with multivariate parameters (Dirichlet, MvNormal, InvWishart) or
deterministic parameters outside the three families raise UnsupportedModel.
"""
import numpy as np

from .. import runtime
from .._lib import (GMM_MU_CONJ, GMM_MU_ESLICE, GMM_MU_HMC, GMM_MU_MWG, GMM_MU_NUTS, GMM_MU_RSLICE, HMC_BLOCK_B,
                    HMC_BLOCK_B_THETA, HMC_BLOCK_THETA, HMC_BLOCK_V, MWG_BLOCK_B, MWG_BLOCK_V)
from . import augurdistlib as adl
from .rvparse import Kern


class UnsupportedModel(NameError):
    """Model or schedule outside the families this runtime executes."""


# ---- expression helpers ----------------------------------------------------

def _var(e):
    return e[1] if e[0] == "var" else None


def _is_lit(e, val=None):
    return e[0] == "lit" and (val is None or float(e[1]) == float(val))


def _proj(e):
    """('proj', ('var', b), idx) -> (b, idx) else (None, None)."""
    if e[0] == "proj" and e[1][0] == "var":
        return e[1][1], e[2]
    return None, None


def _dist(e, name=None):
    if e[0] == "dist" and (name is None or e[1] == name):
        return e[2]
    return None


def _fail(msg):
    raise UnsupportedModel("Error: " + msg)


def _hyper_value(name_or_lit, hypers):
    if name_or_lit[0] == "lit":
        return float(name_or_lit[1])
    n = _var(name_or_lit)
    if n is None or n not in hypers:
        _fail(f"expected a literal or hyper-parameter, found {name_or_lit}")
    return hypers[n]


def _default_sched(model, kinds):
    return [Kern(k, [v]) for v, k in kinds]


def _obj_array(items):
    out = np.empty(len(items), dtype=object)
    for i, it in enumerate(items):
        out[i] = it
    return out


def _ragged(a, dtype):
    """Ragged Vec (Vec t) input (object array, list of arrays or 2-D array)
    -> list of 1-D arrays."""
    if isinstance(a, np.ndarray) and a.dtype != object and a.ndim == 2:
        return [np.asarray(r, dtype=dtype) for r in a]
    return [np.asarray(r, dtype=dtype).ravel() for r in a]


# ---- LDA ---------------------------------------------------------------------

class LdaPlan:
    family = "lda"

    def __init__(self, model, sched):
        data = model.data
        if len(data) != 1 or len(model.params) != 3:
            _fail("not an LDA model")
        w = data[0]
        args = _dist(w.exp, "Categorical")
        if args is None or len(w.idxs) != 2:
            _fail("LDA: data must be w[d, n] ~ Categorical(phi[z[d, n]])")
        phi, pidx = _proj(args[0])
        zname, zidx = _proj(pidx[0]) if pidx and len(pidx) == 1 else (None, None)
        if phi is None or zname is None:
            _fail("LDA: data must be w[d, n] ~ Categorical(phi[z[d, n]])")
        zd = model.decl(zname)
        zargs = _dist(zd.exp, "Categorical") if zd else None
        theta, _ = _proj(zargs[0]) if zargs else (None, None)
        td, pd = model.decl(theta), model.decl(phi)
        if theta is None or td is None or pd is None or len(zd.idxs) != 2:
            _fail("LDA: z[d, n] ~ Categorical(theta[d]) expected")
        ta, pa = _dist(td.exp, "Dirichlet"), _dist(pd.exp, "Dirichlet")
        if ta is None or pa is None:
            _fail("LDA: theta and phi must be Dirichlet")
        self.names = {"theta": theta, "phi": phi, "z": zname, "w": w.name}
        self.alpha_name, self.beta_name = _var(ta[0]), _var(pa[0])
        self.K_name = _var(pd.grids[0][2]) if pd.grids else None
        self.D_name = _var(td.grids[0][2]) if td.grids else None
        self.param_names = [d.name for d in model.params]
        want = [("ConjGibbs", theta), ("ConjGibbs", phi), ("DiscGibbs", zname)]
        if sched is None:
            sched = _default_sched(model, [(v, k) for k, v in want])
        got = [(k.kind, k.vars[0]) for k in sched]
        if len(got) != 3 or got[2] != ("DiscGibbs", zname) or set(got[:2]) != set(want[:2]):
            _fail(f"LDA: supported schedule is ConjGibbs [{theta}] (*) ConjGibbs [{phi}] "
                  f"(*) DiscGibbs [{zname}]")

    def build(self, hypers, data, opt):
        w_docs = _ragged(data[self.names["w"]], np.int32)
        alpha = np.asarray(hypers[self.alpha_name], np.float64)
        beta = np.asarray(hypers[self.beta_name], np.float64)
        K = int(hypers[self.K_name]) if self.K_name in hypers else len(alpha)
        lens = np.array([len(r) for r in w_docs], np.int64)
        self.doc_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        self.D, self.K, self.V = len(w_docs), K, len(beta)
        if len(alpha) != K:
            _fail("LDA: alpha must have K entries")
        w = np.concatenate(w_docs) if w_docs else np.zeros(0, np.int32)
        self.sampler = runtime.LdaSampler(K, self.doc_off, w, alpha, beta, chains=opt.chains,
                                          seed=opt.seed)
        self.alpha, self.beta = alpha, beta
        return self.sampler

    def init_point(self, strat, rng):
        th = _obj_array([adl.initDirichlet(strat, self.alpha, rng) for _ in range(self.D)])
        ph = _obj_array([adl.initDirichlet(strat, self.beta, rng) for _ in range(self.K)])
        z = _obj_array([adl.initCategorical(strat, th[d], rng,
                                            size=int(self.doc_off[d + 1] - self.doc_off[d]))
                        for d in range(self.D)])
        n = self.names
        return {n["theta"]: th, n["phi"]: ph, n["z"]: z}

    def set_point(self, pt, chain):
        n, s = self.names, self.sampler
        if n["theta"] in pt:
            s.set("theta", np.stack([np.asarray(r, np.float64) for r in pt[n["theta"]]]), chain)
        if n["phi"] in pt:
            s.set("phi", np.stack([np.asarray(r, np.float64) for r in pt[n["phi"]]]), chain)
        if n["z"] in pt:
            s.set("z", np.concatenate([np.asarray(r, np.int32).ravel() for r in pt[n["z"]]]),
                  chain)

    def get_point(self, chain):
        n, s = self.names, self.sampler
        th, ph, z = s.get("theta", chain), s.get("phi", chain), s.get("z", chain)
        off = self.doc_off
        return {n["theta"]: _obj_array(list(th)), n["phi"]: _obj_array(list(ph)),
                n["z"]: _obj_array([z[off[d]:off[d + 1]] for d in range(self.D)])}


# ---- HLR ---------------------------------------------------------------------

class HlrPlan:
    family = "hlr"

    def __init__(self, model, sched):
        data = model.data
        if len(data) != 1:
            _fail("not an HLR model")
        yd = data[0]
        bargs = _dist(yd.exp, "Bernoulli")
        self.det_name = None
        if bargs is not None:
            # a deterministic parameter as the Bernoulli bias (dethlr.rv):
            # param bias[n] : Real = sigmoid(...) for n, y[n] ~ Bernoulli(bias[n])
            dn, _ = _proj(bargs[0])
            dd = model.decl(dn) if dn else None
            if dd is not None and dd.kind == "param" and not dd.stochastic:
                self.det_name = dn
                bargs = [dd.exp]
        if bargs is None or bargs[0][0] != "prim" or bargs[0][1] != "sigmoid":
            _fail("HLR: data must be y[n] ~ Bernoulli(sigmoid(...))")
        lin = bargs[0][2][0]
        b_name = None
        if lin[0] == "bin" and lin[1] == "+":
            l, r = lin[2], lin[3]
            if _var(r) and r[0] == "var":
                lin, b_name = l, _var(r)
            elif _var(l):
                lin, b_name = r, _var(l)
        if lin[0] != "prim" or lin[1] != "dotprod" or len(lin[2]) != 2:
            _fail("HLR: linear predictor must be dotprod(theta, x[n]) [+ b]")
        th_name = _var(lin[2][0])
        x_name, xidx = _proj(lin[2][1])
        if th_name is None or x_name is None:
            _fail("HLR: linear predictor must be dotprod(theta, x[n]) [+ b]")
        td = model.decl(th_name)
        targs = _dist(td.exp, "Normal") if td else None
        if targs is None or not _is_lit(targs[0], 0.0):
            _fail("HLR: theta[k] ~ Normal(0.0, v) expected")
        v_name = _var(targs[1])
        vd = model.decl(v_name) if v_name else None
        if vd is None or vd.kind != "param":
            _fail("HLR: the prior variance must be a parameter")
        if _dist(vd.exp, "InvGamma") is not None:
            self.prior = ("InvGamma", _dist(vd.exp, "InvGamma"))
        elif _dist(vd.exp, "Exponential") is not None:
            self.prior = ("Exponential", _dist(vd.exp, "Exponential"))
        else:
            _fail("HLR: v ~ InvGamma(a, b) or v ~ Exponential(lam) expected")
        if b_name is not None:
            bd = model.decl(b_name)
            bargs2 = _dist(bd.exp, "Normal") if bd else None
            if bargs2 is None or not _is_lit(bargs2[0], 0.0) or _var(bargs2[1]) != v_name:
                _fail("HLR: intercept b ~ Normal(0.0, v) expected")
        expected = {v_name, th_name} | ({b_name} if b_name else set())
        stoch = {d.name for d in model.params if d.stochastic}
        dets = {d.name for d in model.params if not d.stochastic}
        if stoch != expected or dets != ({self.det_name} if self.det_name else set()):
            _fail("HLR: unexpected parameters (extra or deterministic parameters)")
        self.names = {"v": v_name, "theta": th_name, "b": b_name, "x": x_name, "y": yd.name,
                      "bias": self.det_name}
        self.param_names = [d.name for d in model.params]
        if sched is None:
            sched = [Kern("HMC", [d.name]) for d in model.params if d.stochastic]
        self.blocks = []
        for k in sched:
            if k.kind == "MWG" and (k.vars == [v_name] or (b_name and k.vars == [b_name])):
                # MWG [v] ~ Normal(v, s) / MWG [b] ~ Normal(b, s): random walks,
                # s the proposal variance (examples/hlr.py:57, test/hlr2.infer)
                nm = k.vars[0]
                prop = k.proposal[0] if k.proposal else None
                pargs = _dist(prop, "Normal") if prop else None
                if pargs is None or _var(pargs[0]) != nm or pargs[1][0] != "lit":
                    _fail(f"HLR: MWG [{nm}] needs the proposal Normal({nm}, <variance>)")
                self.blocks.append((MWG_BLOCK_V if nm == v_name else MWG_BLOCK_B, 0.0,
                                    float(pargs[1][1])))
                continue
            if k.kind != "HMC":
                _fail(f"HLR: only HMC kernels and MWG [{v_name}] run on this runtime (got {k.kind})")
            vs = set(k.vars)
            if vs == {v_name}:
                kind = HMC_BLOCK_V
            elif b_name and vs == {b_name}:
                kind = HMC_BLOCK_B
            elif vs == {th_name}:
                kind = HMC_BLOCK_THETA
            elif b_name and vs == {b_name, th_name}:
                kind = HMC_BLOCK_B_THETA
            else:
                _fail(f"HLR: unsupported HMC block {k.vars}")
            sim_len, eps = k.settings if k.settings else (1.0, 0.05)  # Core/KernSyn.hs:64-75
            self.blocks.append((kind, float(sim_len), float(eps)))

    def build(self, hypers, data, opt):
        n = self.names
        x = np.asarray(np.stack([np.asarray(r, np.float64) for r in hypers[n["x"]]])
                       if np.asarray(hypers[n["x"]]).dtype == object else hypers[n["x"]],
                       np.float64)
        y = np.asarray(data[n["y"]], np.int32).ravel()
        kind, args = self.prior
        if kind == "InvGamma":
            pk, pa, pb = 0, _hyper_value(args[0], hypers), _hyper_value(args[1], hypers)
        else:
            pk, pa, pb = 1, _hyper_value(args[0], hypers), 0.0
        self.prior_vals = (pk, pa, pb)
        blocks = []
        for kind_b, sim_len, eps in self.blocks:
            fixed = 0 if opt.hmc_jitter else max(1, int(round(sim_len / eps)))
            blocks.append((kind_b, sim_len, eps, fixed))
        self.D = x.shape[1]
        self.sampler = runtime.HlrSampler(x.astype(np.float32), y, blocks, chains=opt.chains,
                                          seed=opt.seed, prior_kind=pk, prior_a=pa, prior_b=pb,
                                          has_intercept=n["b"] is not None,
                                          ref_compat=opt.ref_compat)
        return self.sampler

    def init_point(self, strat, rng):
        pk, pa, pb = self.prior_vals
        v = adl.initInverseGamma(strat, pa, pb, rng) if pk == 0 else adl.initExponential(strat, pa, rng)
        n = self.names
        pt = {n["v"]: float(v),
              n["theta"]: np.asarray(adl.initNormal(strat, 0.0, v, rng, size=self.D), np.float64)}
        if n["b"]:
            pt[n["b"]] = float(adl.initNormal(strat, 0.0, v, rng))
        return pt

    def set_point(self, pt, chain):
        n, s = self.names, self.sampler
        cur = s.get("state", max(chain, 0))
        if n["theta"] in pt:
            cur[:self.D] = np.asarray(pt[n["theta"]], np.float64)
        if n["b"] and n["b"] in pt:
            cur[self.D] = float(pt[n["b"]])
        if n["v"] in pt:
            cur[self.D + 1] = float(pt[n["v"]])
        s.set("state", cur, chain)

    def get_point(self, chain):
        n = self.names
        st = self.sampler.get("state", chain)
        pt = {n["v"]: float(st[self.D + 1]), n["theta"]: st[:self.D].copy()}
        if n["b"]:
            pt[n["b"]] = float(st[self.D])
        if n["bias"]:
            pt[n["bias"]] = self.sampler.get("bias", chain)
        return pt


# ---- GMM ---------------------------------------------------------------------

class GmmPlan:
    family = "gmm"

    def __init__(self, model, sched):
        data = model.data
        if len(data) != 1:
            _fail("not a GMM model")
        yd = data[0]
        args = _dist(yd.exp, "MvNormal")
        if args is None or len(args) != 2:
            _fail("GMM: y[n] ~ MvNormal(mu[z[n]], cov[z[n]] | covs) expected")
        mu_name, midx = _proj(args[0])
        z_name, _ = _proj(midx[0]) if midx else (None, None)
        cov_name, cidx = _proj(args[1])
        covs_hyper = None
        if cov_name is None:
            covs_hyper = _var(args[1])
        if mu_name is None or z_name is None or (cov_name is None and covs_hyper is None):
            _fail("GMM: y[n] ~ MvNormal(mu[z[n]], cov[z[n]]) expected")
        zd, md = model.decl(z_name), model.decl(mu_name)
        zargs = _dist(zd.exp, "Categorical") if zd else None
        margs = _dist(md.exp, "MvNormal") if md else None
        if zargs is None or margs is None:
            _fail("GMM: z[n] ~ Categorical(pis), mu[k] ~ MvNormal(mu0, cov0) expected")
        pis_name = _var(zargs[0])
        pd = model.decl(pis_name)
        self.sample_pi = pd is not None and pd.kind == "param"
        if self.sample_pi:
            pa = _dist(pd.exp, "Dirichlet")
            if pa is None:
                _fail("GMM: pis ~ Dirichlet(alpha) expected")
            self.alpha_name = _var(pa[0])
        self.sample_cov = cov_name is not None
        if self.sample_cov:
            cd = model.decl(cov_name)
            ca = _dist(cd.exp, "InvWishart") if cd else None
            if ca is None:
                _fail("GMM: cov[k] ~ InvWishart(df, scale0) expected")
            self.df_name, self.scale0_name = _var(ca[0]), _var(ca[1])
        self.mu0_name, self.cov0_name = _var(margs[0]), _var(margs[1])
        self.names = {"pis": pis_name, "mu": mu_name, "cov": cov_name or covs_hyper, "z": z_name,
                      "y": yd.name}
        self.param_names = [d.name for d in model.params]
        want = []
        if self.sample_pi:
            want.append(("ConjGibbs", pis_name))
        want.append(("ConjGibbs", mu_name))
        if self.sample_cov:
            want.append(("ConjGibbs", cov_name))
        want.append(("DiscGibbs", z_name))
        if sched is None:
            sched = [Kern(k, [v]) for k, v in want]
        # the kernel for mu may be ConjGibbs, ESlice, HMC, MWG or RSlice
        # (compiler/augur/test/hmvgmm1-3.infer, mvgmm1-5.infer)
        mu_kerns = {"ConjGibbs", "ESlice", "HMC", "MWG", "RSlice", "NUTS"}
        got = [(k.kind, k.vars[0]) if len(k.vars) == 1 else (k.kind, tuple(k.vars)) for k in sched]
        norm = [("ConjGibbs", v) if (v == mu_name and kd in mu_kerns) else (kd, v) for kd, v in got]
        if norm != want:
            _fail("GMM: supported schedule is " + " (*) ".join(
                f"{'ConjGibbs|ESlice|HMC|MWG|RSlice|NUTS' if v == mu_name else k} [{v}]" for k, v in want))
        mk = next(k for k in sched if k.vars == [mu_name])
        self.mu_kernel = {"ConjGibbs": GMM_MU_CONJ, "ESlice": GMM_MU_ESLICE, "HMC": GMM_MU_HMC,
                          "MWG": GMM_MU_MWG, "RSlice": GMM_MU_RSLICE, "NUTS": GMM_MU_NUTS}[mk.kind]
        # Core/KernSyn.hs:64-75, :94-117: HMC and RSlice default to (1.0, 0.05),
        # NUTS to the step size 0.05
        if mk.kind == "NUTS":
            self.mu_settings = (0.0, float(mk.settings[0]) if mk.settings else 0.05)
        else:
            self.mu_settings = tuple(mk.settings) if mk.settings else (1.0, 0.05)
        self.mu_prop_cov_name = None
        if mk.kind == "MWG":
            # MWG [mu] ~ MvNormal(mu[k], P) for k <- 0 until K
            prop, grids = mk.proposal
            pargs = _dist(prop, "MvNormal")
            pname, pidx = _proj(pargs[0]) if pargs else (None, None)
            if (pargs is None or pname != mu_name or len(grids) != 1 or not pidx
                    or _var(pidx[0]) != grids[0][0] or _var(pargs[1]) is None):
                _fail(f"GMM: MWG [{mu_name}] needs the proposal MvNormal({mu_name}[k], <cov hyper>) "
                      f"for k <- 0 until K")
            self.mu_prop_cov_name = _var(pargs[1])

    def build(self, hypers, data, opt):
        y = np.asarray(np.stack([np.asarray(r, np.float64) for r in data[self.names["y"]]]),
                       np.float64)
        self.N, self.Dm = y.shape
        mu0 = np.asarray(hypers[self.mu0_name], np.float64)
        cov0 = np.asarray(hypers[self.cov0_name], np.float64)
        if self.sample_pi:
            alpha = np.asarray(hypers[self.alpha_name], np.float64)
            self.K = len(alpha)
        else:
            pis = np.asarray(hypers[self.names["pis"]], np.float64)
            self.K = len(pis)
            alpha = np.ones(self.K)
        if self.sample_cov:
            df = int(hypers[self.df_name])
            scale0 = np.asarray(hypers[self.scale0_name], np.float64)
        else:
            df, scale0 = self.Dm + 2, np.eye(self.Dm)
        self.hyp = dict(alpha=alpha, mu0=mu0, cov0=cov0, df=df, scale0=scale0)
        sim_len, eps = self.mu_settings
        prop_cov = (np.asarray(hypers[self.mu_prop_cov_name], np.float64)
                    if self.mu_prop_cov_name else None)
        fixed = 0 if opt.hmc_jitter else max(1, int(round(sim_len / eps)))
        s = runtime.GmmSampler(y.astype(np.float32), alpha, mu0, cov0, df, scale0,
                               chains=opt.chains, seed=opt.seed, sample_pi=self.sample_pi,
                               sample_cov=self.sample_cov, mu_kernel=self.mu_kernel,
                               mu_sim_len=sim_len, mu_step_size=eps,
                               mu_fixed_steps=fixed if self.mu_kernel == GMM_MU_HMC else 0,
                               mu_prop_cov=prop_cov, ref_compat=opt.ref_compat)
        if not self.sample_pi:
            s.set("pis", pis)
        if not self.sample_cov:
            covs = np.asarray(hypers[self.names["cov"]], np.float64)
            s.set("cov", np.broadcast_to(covs, (self.K, self.Dm, self.Dm)))
        self.sampler = s
        return s

    def init_point(self, strat, rng):
        h, n = self.hyp, self.names
        pt = {}
        pis = adl.initDirichlet(strat, h["alpha"], rng) if self.sample_pi else None
        if self.sample_pi:
            pt[n["pis"]] = pis
        pt[n["mu"]] = _obj_array([adl.initMvNormal(strat, h["mu0"], h["cov0"], rng)
                                  for _ in range(self.K)])
        if self.sample_cov:
            pt[n["cov"]] = _obj_array([adl.initInvWishart(strat, h["df"], h["scale0"], rng)
                                       for _ in range(self.K)])
        pmf = pis if pis is not None else np.full(self.K, 1.0 / self.K)
        pt[n["z"]] = adl.initCategorical(strat, pmf, rng, size=self.N)
        return pt

    def set_point(self, pt, chain):
        n, s = self.names, self.sampler
        if self.sample_pi and n["pis"] in pt:
            s.set("pis", np.asarray(pt[n["pis"]], np.float64), chain)
        if n["mu"] in pt:
            s.set("mu", np.stack([np.asarray(r, np.float64) for r in pt[n["mu"]]]), chain)
        if self.sample_cov and n["cov"] in pt:
            s.set("cov", np.stack([np.asarray(r, np.float64) for r in pt[n["cov"]]]), chain)
        if n["z"] in pt:
            s.set("z", np.asarray(pt[n["z"]], np.int32).ravel(), chain)

    def get_point(self, chain):
        n, s = self.names, self.sampler
        pt = {n["mu"]: _obj_array(list(s.get("mu", chain))), n["z"]: s.get("z", chain)}
        if self.sample_pi:
            pt[n["pis"]] = s.get("pis", chain)
        if self.sample_cov:
            pt[n["cov"]] = _obj_array([np.asarray(m) for m in s.get("cov", chain)])
        return pt


def _np_eval(e, env):
    """Value of a prior argument for initialisation (numpy, vectorised over
    the grid variables bound in env)."""
    k = e[0]
    if k == "lit":
        return e[1]
    if k == "var":
        return env[e[1]]
    if k == "proj":
        base = np.asarray(_np_eval(e[1], env))
        idx = tuple(np.asarray(_np_eval(i, env)).astype(np.int64) for i in e[2])
        return base[idx]
    if k == "bin":
        a, b = _np_eval(e[2], env), _np_eval(e[3], env)
        return {"+": lambda: a + b, "-": lambda: a - b, "*": lambda: a * b, "/": lambda: a / b}[e[1]]()
    if k == "prim" and e[1] == "sigmoid":
        return 1.0 / (1.0 + np.exp(-_np_eval(e[2][0], env)))
    if k == "prim" and e[1] == "dotprod":
        return np.sum(np.asarray(_np_eval(e[2][0], env)) * np.asarray(_np_eval(e[2][1], env)), axis=-1)
    _fail(f"cannot evaluate {e} for initialisation")


class GenPlan:
    """Any model whose parameters are continuous or finite discrete: the
    passes are generated from the model (pyaugur.lower) and run as HMC
    blocks on continuous parameters and DiscGibbs on discrete ones.  The
    reference's kernel heuristic gives each parameter its own kernel:
    DiscGibbs for a discrete one, else HMC with the defaults simLen 1.0,
    eps 0.05 (Compile/Compile.hs:143-165, Core/KernSyn.hs:64-75); it would
    pick ConjGibbs where it detects conjugacy, which the generic path does
    not analyse (HMC targets the same posterior)."""

    family = "gen"

    def __init__(self, model, sched):
        from .lower import DATA_DISTS, DISC_DISTS, PARAM_DISTS
        for d in model.params:
            if not d.stochastic or d.exp[0] != "dist" or d.exp[1] not in PARAM_DISTS | DISC_DISTS:
                _fail(f"generic: parameter {d.name} must be continuous "
                      f"({', '.join(sorted(PARAM_DISTS))}) or discrete ({', '.join(sorted(DISC_DISTS))})")
        for d in model.data:
            if d.exp[0] != "dist" or d.exp[1] not in DATA_DISTS:
                _fail(f"generic: data {d.name} must follow one of {', '.join(sorted(DATA_DISTS))}")
        self.model = model
        self.param_names = [d.name for d in model.params]
        self.discrete = {d.name for d in model.params if d.exp[1] in DISC_DISTS}
        if sched is None:
            sched = [Kern("DiscGibbs" if n in self.discrete else "HMC", [n]) for n in self.param_names]
        self.blocks = []
        seen = set()
        for k in sched:
            if k.kind not in ("HMC", "DiscGibbs"):
                _fail(f"generic: the generic path runs HMC and DiscGibbs blocks (got {k.kind})")
            for v in k.vars:
                if v not in self.param_names:
                    _fail(f"generic: {v} is not a parameter")
                if v in seen:
                    _fail(f"generic: {v} appears in two blocks")
                if (k.kind == "DiscGibbs") != (v in self.discrete):
                    _fail(f"generic: {v} needs {'DiscGibbs' if v in self.discrete else 'HMC'}")
                seen.add(v)
            if k.kind == "DiscGibbs":
                self.blocks.append(("DiscGibbs", list(k.vars)))
                continue
            sim_len, eps = k.settings if k.settings else (1.0, 0.05)
            self.blocks.append((list(k.vars), float(sim_len), float(eps)))
        missing = [n for n in self.param_names if n not in seen]
        if missing:
            _fail(f"generic: no kernel updates {', '.join(missing)}")

    def build(self, hypers, data, opt):
        from .lower import Lowering, LoweringError
        lblocks = [b if b[0] == "DiscGibbs" else b[0] for b in self.blocks]
        try:
            self.lowering = Lowering(self.model, hypers, data, ref_compat=opt.ref_compat, blocks=lblocks)
        except LoweringError as e:
            _fail(f"generic: {e}")
        self.hypers = hypers
        blocks = []
        for b in self.blocks:
            if b[0] == "DiscGibbs":
                blocks.append(("DiscGibbs", b[1][0]))
                continue
            vars_, sim_len, eps = b
            fixed = 0 if opt.hmc_jitter else max(1, int(round(sim_len / eps)))
            blocks.append((vars_, sim_len, eps, fixed))
        self.sampler = runtime.GenSampler(self.lowering, blocks, chains=opt.chains, seed=opt.seed)
        return self.sampler

    def _shape(self, name):
        lw = self.lowering
        return lw.layout[name][1] if name in lw.layout else lw.zlayout[name][1]

    def init_point(self, strat, rng):
        """Initial values from each parameter's prior, in declaration order
        (augurdistlib's init* per distribution)."""
        env = {k: (np.asarray(v) if not np.isscalar(v) else v) for k, v in self.hypers.items()}
        pt = {}
        for d in self.model.params:
            shape = self._shape(d.name)
            g = dict(env)
            if shape:
                idx = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
                for (var, _, _), ix in zip(d.grids, idx):
                    g[var] = ix
            size = shape if shape else None
            name = d.exp[1]
            if name == "Categorical":
                pmf = np.asarray(_np_eval(d.exp[2][0], g), np.float64)
                v = adl.initCategorical(strat, pmf, rng, size=size)
                pt[d.name] = np.asarray(v, np.int32) if shape else int(v)
                env[d.name] = pt[d.name]
                continue
            args = [np.broadcast_to(np.asarray(_np_eval(a, g), np.float64), shape) if shape
                    else float(_np_eval(a, g)) for a in d.exp[2]]
            if name == "Bernoulli":
                v = (np.zeros(shape, np.int32) if isinstance(strat, adl.AugurDefault) else
                     (rng.random(size) < np.asarray(args[0])).astype(np.int32))
                pt[d.name] = np.asarray(v, np.int32) if shape else int(v)
                env[d.name] = pt[d.name]
                continue
            if name == "Normal":
                v = adl.initNormal(strat, args[0], args[1], rng, size=size)
            elif name == "Exponential":
                v = (np.full(shape, 1.0) if isinstance(strat, adl.AugurDefault) else
                     rng.exponential(1.0 / np.asarray(args[0]), size=size)) if shape else \
                    adl.initExponential(strat, args[0], rng)
            elif name == "InvGamma":
                v = (np.full(shape, 0.5) if isinstance(strat, adl.AugurDefault) else
                     1.0 / rng.gamma(args[0], 1.0 / np.asarray(args[1]), size=size)) if shape else \
                    adl.initInverseGamma(strat, args[0], args[1], rng)
            elif name == "Gamma":
                v = adl.initGamma(strat, args[0], args[1], rng, size=size)
            elif name == "Beta":
                v = adl.initBeta(strat, args[0], args[1], rng, size=size)
            else:
                v = adl.initUniform(strat, args[0], args[1], rng, size=size)
            pt[d.name] = np.asarray(v, np.float64) if shape else float(v)
            env[d.name] = pt[d.name]
        return pt

    def set_point(self, pt, chain):
        for name in self.param_names:
            if name in pt:
                dt = np.int32 if name in self.discrete else np.float64
                self.sampler.set(name, np.asarray(pt[name], dt), chain)

    def get_point(self, chain):
        out = {}
        for name in self.param_names:
            v = self.sampler.get(name, chain)
            if not self._shape(name):
                out[name] = int(v[0]) if name in self.discrete else float(v[0])
            else:
                out[name] = v
        return out


def match(model, sched, generic=False):
    """Return the plan of the first family that matches, else the generic
    lowering's plan, else raise.  generic=True skips the families."""
    errors = []
    for cls in (() if generic else (LdaPlan, HlrPlan, GmmPlan)) + (GenPlan,):
        try:
            return cls(model, sched)
        except UnsupportedModel as e:
            errors.append(f"{cls.family}: {e}")
        except (AttributeError, IndexError, TypeError) as e:
            errors.append(f"{cls.family}: {e!r}")
    raise UnsupportedModel("Error: the model is outside what this runtime executes on the GPU "
                           "(LDA, hierarchical logistic regression, Gaussian mixture, or a model "
                           "with continuous parameters only): " + "; ".join(errors))
