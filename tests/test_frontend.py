"""Host logic without a GPU: parser, plan matching, the C-ABI library's
exported symbols, error behaviour, and the diagnostics."""
import ctypes
import os
import re

import numpy as np
import pytest

from augurv2_amd.pyaugur import plan, rvparse
from augurv2_amd.pyaugur.augurlib import AugurInfer, AugurOpt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

LDA = """(K : Int, D : Int, N : Vec Int, alpha : Vec Real, beta : Vec Real) => {
  param theta[d] ~ Dirichlet(alpha)
      for d <- 0 until D ;
  param phi[k] ~ Dirichlet(beta)
      for k <- 0 until K ;
  param z[d, n] ~ Categorical(theta[d])
      for d <- 0 until D, n <- 0 until N[d] ;
  data w[d, n] ~ Categorical(phi[z[d, n]])
      for d <- 0 until D, n <- 0 until N[d] ;
}"""

HLR_CFG = """(K : Int, N : Int, x : Vec (Vec Real)) => {
  param v ~ InvGamma(1.0, 1.0) ;
  param beta[k] ~ Normal(0.0, v) for k <- 0 until K ;
  data y[n] ~ Bernoulli(sigmoid(dotprod(beta, x[n]))) for n <- 0 until N ;
}"""

HLR = """(K : Int, N : Int, lam : Real, x : Vec (Vec Real)) => {
  param v ~ Exponential(lam) ;
  param b ~ Normal(0.0, v) ;
  param theta[k] ~ Normal(0.0, v)
      for k <- 0 until K ;
  data y[n] ~ Bernoulli( sigmoid(dotprod(theta, x[n]) + b) )
      for n <- 0 until N ;
}"""

HMVGMM = """(K : Int, N : Int, alpha : Vec Real, mu0 : Vec Real, covs0 : Mat Real, df : Int, scale0 : Mat Real) => {
  param pis ~ Dirichlet(alpha) ;
  param mu[k] ~ MvNormal(mu0, covs0)
      for k <- 0 until K ;
  param cov[k] ~ InvWishart(df, scale0)
      for k <- 0 until K ;
  param z[n] ~ Categorical(pis)
      for n <- 0 until N ;
  data y[n] ~ MvNormal(mu[z[n]], cov[z[n]])
      for n <- 0 until N ;
}"""


def test_parse_lda_structure():
    m = rvparse.parse_model(LDA)
    assert [h[0] for h in m.hypers] == ["K", "D", "N", "alpha", "beta"]
    assert m.hypers[2][1] == ("Vec", ("Int",))
    z = m.decl("z")
    assert z.idxs == ["d", "n"] and z.stochastic
    assert z.grids[1][2] == ("proj", ("var", "N"), [("var", "d")])
    assert m.decl("w").exp == ("dist", "Categorical",
                               [("proj", ("var", "phi"),
                                 [("proj", ("var", "z"), [("var", "d"), ("var", "n")])])])


def test_parse_precedence_and_deterministic():
    m = rvparse.parse_model("(a : Real) => { param p : Real = a + 2.0 * a - 1 ; }")
    e = m.decl("p").exp
    assert e == ("bin", "-", ("bin", "+", ("var", "a"), ("bin", "*", ("lit", 2.0), ("var", "a"))),
                 ("lit", 1))
    assert not m.decl("p").stochastic


def test_parse_schedules():
    k = rvparse.parse_schedule("HMC [v] (*) HMC [b, theta] [0.1, 0.005] (*) DiscGibbs [z]")
    assert [(x.kind, x.vars, x.settings) for x in k] == [
        ("HMC", ["v"], None), ("HMC", ["b", "theta"], (0.1, 0.005)), ("DiscGibbs", ["z"], None)]
    k = rvparse.parse_schedule("MWG [mu] ~ MvNormal(mu[k], covs0) for k <- 0 until K (*) NUTS [mu] [0.05]")
    assert k[0].kind == "MWG" and k[1].settings == (0.05,)
    for bad in ["HMC v", "ConjGibbs [a, b]", "HMC [v] [1.0]", "Foo [x]", "HMC [v] (*)"]:
        with pytest.raises(rvparse.RvParseError):
            rvparse.parse_schedule(bad)


def test_parse_errors():
    for bad in ["(K : Int) => { param x ~ Normal(0.0, 1.0) }",  # missing ';'
                "(K : Int) { }", "(K : Foo) => { }", "(K : Int) => { param x ~ ; }"]:
        with pytest.raises(rvparse.RvParseError):
            rvparse.parse_model(bad)


REF_TEST = "/root/reference/compiler/augur/test"


@pytest.mark.skipif(not os.path.isdir(REF_TEST), reason="reference tree not mounted")
def test_reference_models_and_schedules_parse():
    import glob
    for f in glob.glob(os.path.join(REF_TEST, "*.rv")):
        rvparse.parse_model(open(f).read())
    for f in glob.glob(os.path.join(REF_TEST, "*.infer")):
        rvparse.parse_schedule(open(f).read())


@pytest.mark.skipif(not os.path.isdir(REF_TEST), reason="reference tree not mounted")
def test_reference_schedules_match_a_plan():
    """Every model / schedule pair of compiler/augur/test matches a sampler
    family, except the ones listed as outside this runtime."""
    pairs = {"lda1": "lda", "hlr1": "hlr", "hlr2": "hlr", "hmvgmm1": "hmvgmm", "hmvgmm2": "hmvgmm",
             "hmvgmm3": "hmvgmm", "mvgmm1": "mvgmm", "mvgmm2": "mvgmm", "mvgmm3": "mvgmm",
             "mvgmm4": "mvgmm", "mvgmm5": "mvgmm", "mvgmm6": "mvgmm"}
    unsupported = set()
    dm = rvparse.parse_model(open(os.path.join(REF_TEST, "dethlr.rv")).read())
    assert plan.match(dm, None).names["bias"] == "bias"  # deterministic parameter
    got = {}
    for sched, model in pairs.items():
        m = rvparse.parse_model(open(os.path.join(REF_TEST, model + ".rv")).read())
        k = rvparse.parse_schedule(open(os.path.join(REF_TEST, sched + ".infer")).read())
        try:
            plan.match(m, k)
            got[sched] = True
        except plan.UnsupportedModel:
            got[sched] = False
    assert {s for s, ok in got.items() if not ok} <= unsupported, got
    for s in ("hmvgmm2", "hmvgmm3", "mvgmm2", "mvgmm3", "mvgmm4", "mvgmm5", "mvgmm6", "hlr2"):
        assert got[s], s


def test_plan_matching():
    p = plan.match(rvparse.parse_model(LDA),
                   rvparse.parse_schedule("ConjGibbs [theta] (*) ConjGibbs [phi] (*) DiscGibbs [z]"))
    assert p.family == "lda"
    p = plan.match(rvparse.parse_model(LDA), None)  # automatic schedule
    assert p.family == "lda"
    p = plan.match(rvparse.parse_model(HLR_CFG), rvparse.parse_schedule("HMC [v] (*) HMC [beta] [0.1, 0.005]"))
    assert p.family == "hlr" and p.blocks == [(0, 1.0, 0.05), (2, 0.1, 0.005)]
    p = plan.match(rvparse.parse_model(HLR), rvparse.parse_schedule("HMC [v] (*) HMC [b, theta]"))
    assert p.blocks[1][0] == 3 and p.names["b"] == "b"
    p = plan.match(rvparse.parse_model(HMVGMM), None)
    assert p.family == "gmm" and p.sample_pi and p.sample_cov
    # examples/hlr.py:57's random-walk MWG on the variance
    p = plan.match(rvparse.parse_model(HLR_CFG),
                   rvparse.parse_schedule("MWG [v] ~ Normal(v, 1.0) (*) HMC [beta] [0.1, 0.005]"))
    assert p.blocks == [(4, 0.0, 1.0), (2, 0.1, 0.005)]
    p = plan.match(rvparse.parse_model(HLR), rvparse.parse_schedule(  # test/hlr2.infer
        "MWG [v] ~ Normal(v, 1.0) (*) MWG [b] ~ Normal(b, 1.0) (*) HMC [theta]"))
    assert p.blocks == [(4, 0.0, 1.0), (5, 0.0, 1.0), (2, 1.0, 0.05)]
    with pytest.raises(plan.UnsupportedModel):  # only the random walk centred at v
        plan.match(rvparse.parse_model(HLR_CFG), rvparse.parse_schedule(
            "MWG [v] ~ Normal(1.0, 1.0) (*) HMC [beta]"))
    p = plan.match(rvparse.parse_model(HMVGMM), rvparse.parse_schedule(
        "ConjGibbs [pis] (*) ESlice [mu] (*) ConjGibbs [cov] (*) DiscGibbs [z]"))
    assert p.mu_kernel == 1
    p = plan.match(rvparse.parse_model(HMVGMM), rvparse.parse_schedule(
        "ConjGibbs [pis] (*) HMC [mu] [0.5, 0.05] (*) ConjGibbs [cov] (*) DiscGibbs [z]"))
    assert p.mu_kernel == 2 and p.mu_settings == (0.5, 0.05)
    with pytest.raises(plan.UnsupportedModel):  # the kernels must keep the order pis, mu, cov, z
        plan.match(rvparse.parse_model(HMVGMM), rvparse.parse_schedule(
            "ConjGibbs [pis] (*) ConjGibbs [cov] (*) ESlice [mu] (*) DiscGibbs [z]"))
    with pytest.raises(NameError):  # UnsupportedModel is the reference's NameError
        plan.match(rvparse.parse_model("(a : Real) => { param p ~ Poisson(a) ; }"), None)
    # a continuous model outside the families goes to the generic lowering,
    # one HMC block per parameter with the defaults (1.0, 0.05)
    p = plan.match(rvparse.parse_model("(a : Real) => { param p ~ Normal(0.0, a) ; }"), None)
    assert p.family == "gen" and p.blocks == [(["p"], 1.0, 0.05)]
    with pytest.raises(NameError):  # the generic path runs HMC blocks only
        plan.match(rvparse.parse_model("(a : Real) => { param p ~ Normal(0.0, a) ; }"),
                   rvparse.parse_schedule("ESlice [p]"))
    p = plan.match(rvparse.parse_model(HLR), None, generic=True)
    assert p.family == "gen" and [b[0] for b in p.blocks] == [["v"], ["b"], ["theta"]]


def test_augur_infer_errors_without_gpu_work():
    with AugurInfer(None, LDA) as infer:
        with pytest.raises(NameError):
            infer.compile(2)()  # no compile options yet
        infer.set_compile_opt(AugurOpt(target="cpu"))
        with pytest.raises(NameError):
            infer.compile(2)()


class _CountingSampler:
    """Stands in for a runtime sampler: counts sweeps."""

    def __init__(self):
        self.sweeps = 0

    def step(self, n=1):
        self.sweeps += n

    def close(self):
        pass


class _CountingPlan:
    param_names = ["x"]

    def __init__(self, sampler):
        self.sampler = sampler

    def get_point(self, chain):
        return {"x": self.sampler.sweeps}


def test_samplen_keeps_the_reference_sweeps():
    # pyaugur/augurlib.py:195-222: numSamples * thin sweeps after burnIn, the
    # state kept after sweep i when i % thin == 0
    for burn, num, thin in ((2, 3, 4), (0, 5, 1), (1, 2, 3)):
        with AugurInfer(None, LDA) as infer:
            infer.aopt = AugurOpt(target="gpu")
            infer.sampler = _CountingSampler()
            infer.plan = _CountingPlan(infer.sampler)
            infer.param_names = ["x"]
            out = infer.samplen(burnIn=burn, numSamples=num, thin=thin)
            ref = [burn + i + 1 for i in range(num * thin) if i % thin == 0]
            assert out["x"] == ref
            assert infer.sampler.sweeps == burn + num * thin
            infer.sampler = None


def test_augur_opt_defaults_match_the_reference():
    # pyaugur/augurlib.py:39: AugurOpt(cached=False, target='cpu', initStrat=AugurRandom(), paramScale=None)
    opt = AugurOpt()
    assert opt.target == "cpu" and opt.cached is False and opt.paramScale is None
    assert type(opt.initStrat).__name__ == "AugurRandom"


# ---- the C-ABI library loads and exports every declared symbol ----

def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "augur_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(augur_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_declared_symbols():
    from augurv2_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared_symbols()
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES)


def test_library_fails_loudly_without_device():
    from augurv2_amd import _lib
    if _lib.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(_lib.AugurError):
        _lib.require_gpu()
    # a compute entry point reports the missing device instead of computing
    from augurv2_amd import ops
    with pytest.raises(_lib.AugurError, match="no HIP device"):
        ops.gamma(0, 0, 0, 1, np.ones(3))


def test_library_argument_errors():
    from augurv2_amd import _lib
    L = _lib.load()
    assert L.augur_destroy(None) == 0
    assert L.augur_step(None, 1) == 1
    assert b"null sampler" in L.augur_last_error()
    assert L.augur_lda_create(None, None) == 1


# ---- diagnostics ----

def _ar1(rng, rho, chains, n):
    x = np.zeros((chains, n))
    e = rng.standard_normal((chains, n))
    x[:, 0] = e[:, 0] / np.sqrt(1 - rho**2)
    for i in range(1, n):
        x[:, i] = rho * x[:, i - 1] + e[:, i]
    return x


def test_ess_matches_ar1_closed_form():
    from augurv2_amd import diagnostics
    rng = np.random.default_rng(0)
    for rho in (0.0, 0.5, 0.9):
        ests = [diagnostics.ess_bulk(_ar1(rng, rho, 4, 8000)) for _ in range(5)]
        expect = 4 * 8000 * (1 - rho) / (1 + rho)
        assert abs(np.mean(ests) / expect - 1) < 0.12, (rho, np.mean(ests), expect)


def test_rhat_detects_shifted_chain():
    from augurv2_amd import diagnostics
    rng = np.random.default_rng(1)
    x = rng.standard_normal((4, 2000))
    assert diagnostics.rhat(x) < 1.01
    x[0] += 1.0
    assert diagnostics.rhat(x) > 1.1
    # a chain stuck at a different scale is caught by the folded (tail) R-hat
    y = rng.standard_normal((4, 2000))
    y[1] *= 3
    assert diagnostics.rhat(y) > 1.05


def test_mcse_iid():
    from augurv2_amd import diagnostics
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 5000))
    assert abs(diagnostics.mcse_mean(x) / (1 / np.sqrt(20000)) - 1) < 0.1
