"""The generic lowering on the GPU: the generated data pass (compiled with
hipRTC) against the oracle's log joint and gradient, HMC trajectories
against the oracle's restatement, and posteriors against closed forms and
the hand-written HLR sampler."""
import math

import numpy as np
import pytest

from augurv2_amd import runtime
from augurv2_amd.pyaugur.augurlib import AugurInfer, AugurOpt, AugurUser
from augurv2_amd.pyaugur.lower import Lowering
from augurv2_amd.pyaugur.rvparse import parse_model
from oracle import genmodel as gm
from tests.gen_models import HLR, models, point

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    from augurv2_amd import _lib
    _lib.require_gpu()


def _pair_data(name, n, seed=7):
    _, hyp, dat = models(np.random.default_rng(seed), n)[name]
    return hyp, dat


def _pair(name, n, ref_compat, seed=7):
    rng = np.random.default_rng(seed)
    src, hyp, dat = models(rng, n)[name]
    model = parse_model(src)
    return Lowering(model, hyp, dat, ref_compat), gm.GenModel(model, hyp, dat, ref_compat)


@pytest.mark.parametrize("ref_compat", [True, False])
@pytest.mark.parametrize("name,n", [("linreg", 3000), ("hlr", 5000), ("pois", 4000), ("beta", 2500),
                                    ("mixed", 3000), ("hlr", 130_000)])
def test_pass_matches_oracle(name, n, ref_compat):
    """ll and d ll / dq of the generated pass at random points of 3 chains
    (FP64 on both sides; only the summation order differs)."""
    print(f"generic pass {name} n={n} ref_compat={ref_compat}", flush=True)
    lw, om = _pair(name, n, ref_compat)
    rng = np.random.default_rng(11)
    q = np.stack([point(om, rng) for _ in range(3)])
    with runtime.GenSampler(lw, [], chains=3, seed=1) as s:
        ll, g = s.eval(q)
    for c in range(3):
        assert ll[c] == pytest.approx(om.ll(q[c]), rel=1e-11, abs=1e-9)
        ref = om.grad(q[c])
        np.testing.assert_allclose(g[c], ref, rtol=1e-9, atol=1e-9 * max(1.0, np.abs(ref).max()))


def test_pass_grid_covers_every_row():
    # more rows than the largest grid holds at four per thread: the
    # grid-stride loop takes the rest; a dropped or doubled row changes ll
    lw, om = _pair("hlr", 600_001, False)
    assert lw.grid == 512
    rng = np.random.default_rng(12)
    q = point(om, rng)[None]
    with runtime.GenSampler(lw, [], chains=1, seed=1) as s:
        ll, _ = s.eval(q)
    assert ll[0] == pytest.approx(om.ll(q[0]), rel=1e-11)


@pytest.mark.parametrize("name,blocks", [
    ("hlr", [(["v"], 1.0, 0.05), (["b", "theta"], 0.5, 0.05)]),
    ("linreg", [(["s2"], 0.5, 0.1), (["w"], 0.5, 0.05)]),
    # the reference's variance gradient (normal_dotarg2's -0.5 var) is not
    # the derivative of its normal_ll, so a block holding tau rejects almost
    # every trajectory; the w block accepts
    ("mixed", [(["w"], 0.3, 0.03), (["tau"], 0.1, 0.01)]),
])
def test_hmc_trajectories_match_oracle(name, blocks):
    """Three sweeps from the same point with the same Philox streams: every
    chain's state after each sweep equals the oracle's restatement of
    cbits/augur_mcmclib.h:343-424 (same jitter, momenta and decisions)."""
    print(f"generic HMC {name}", flush=True)
    lw, om = _pair(name, 400, True)
    # each block runs the pass of its own full conditional
    lw = Lowering(lw.model, *_pair_data(name, 400), ref_compat=True, blocks=[b[0] for b in blocks])
    assert len(lw.block_kernels) == len(blocks)
    rng = np.random.default_rng(13)
    chains, seed = 2, 21
    q0 = np.stack([point(om, rng) for _ in range(chains)])
    with runtime.GenSampler(lw, [(v, sl, ep, 0) for v, sl, ep in blocks], chains=chains, seed=seed) as s:
        s.set("state", q0[0], 0)
        s.set("state", q0[1], 1)
        cur = q0.copy()
        accepted = 0
        for sweep in range(3):
            s.step(1)
            for c in range(chains):
                for b, (vars_, sl, ep) in enumerate(blocks):
                    mask = np.zeros(om.P, bool)
                    for v in vars_:
                        off, shape = om.layout[v]
                        mask[off:off + (int(np.prod(shape)) if shape else 1)] = True
                    cur[c], acc = om.hmc_step(cur[c], mask, sl, ep, 0, seed, c, sweep, b, block_vars=vars_)
                    accepted += acc
                np.testing.assert_allclose(s.get("state", c), cur[c], rtol=1e-9, atol=1e-11)
        assert accepted > 0
        acc = s.acceptance()
        assert acc.shape == (len(blocks), chains) and np.all((acc >= 0) & (acc <= 1))
        ll = s.log_density()
        for c in range(chains):
            assert ll[c] == pytest.approx(om.ll(cur[c]), rel=1e-11)


def test_conjugate_posterior():
    """mu ~ Normal(m0, v0), y[n] ~ Normal(mu, s2) (textbook densities): the
    HMC draws' mean and variance match the closed-form posterior."""
    src = """(N : Int, m0 : Real, v0 : Real, s2 : Real) => {
      param mu ~ Normal(m0, v0) ;
      data y[n] ~ Normal(mu, s2) for n <- 0 until N ;
    }"""
    rng = np.random.default_rng(14)
    N, s2 = 5000, 4.0
    y = rng.normal(-0.7, 2.0, N)
    V = 1.0 / (1.0 / 10.0 + N / s2)
    mean = (0.5 / 10.0 + y.sum() / s2) * V
    with AugurInfer(None, src) as infer:
        infer.set_compile_opt(AugurOpt(target="gpu", seed=3, chains=4, ref_compat=False,
                                       initStrat=AugurUser({"mu": 0.0})))
        infer.set_user_sched(f"HMC [mu] [{3 * math.sqrt(V)}, {0.3 * math.sqrt(V)}]")
        infer.compile(N, 0.5, 10.0, s2)(y)
        s = infer.samplen(burnIn=200, numSamples=2000)
    d = np.array(s["mu"])  # samples x chains
    assert abs(d.mean() - mean) < 4 * math.sqrt(V / 2000)
    assert 0.85 < d.var() / V < 1.15


def test_hlr_through_the_generic_path_matches_the_family_sampler():
    """hlr.rv lowered generically (FP64 data) and run by the hand-written
    HLR sampler (FP32 data): posterior means of theta and b agree."""
    rng = np.random.default_rng(15)
    K, N = 5, 3000
    x = rng.standard_normal((N, K))
    theta = np.array([1.0, -1.0, 0.5, 0.0, 2.0])
    y = (rng.random(N) < 1 / (1 + np.exp(-(x @ theta - 0.5)))).astype(np.int32)
    out = {}
    for generic in (True, False):
        with AugurInfer(None, HLR) as infer:
            infer.set_compile_opt(AugurOpt(target="gpu", seed=5, chains=4, generic=generic, ref_compat=False,
                                           initStrat=AugurUser({"v": 1.0, "b": 0.0, "theta": np.zeros(K)})))
            infer.set_user_sched("HMC [v] [0.5, 0.05] (*) HMC [b] [0.5, 0.05] (*) HMC [theta] [0.5, 0.05]")
            infer.compile(K, N, 1.0, x)(y)
            assert infer.plan.family == ("gen" if generic else "hlr")
            s = infer.samplen(burnIn=300, numSamples=1500)
        out[generic] = (np.array(s["theta"]).reshape(-1, K), np.array(s["b"]).ravel())
    th_g, b_g = out[True]
    th_h, b_h = out[False]
    # posterior sd of each coefficient is about 0.05; 6000 draws per sampler
    # with autocorrelation: allow 0.04
    np.testing.assert_allclose(th_g.mean(0), th_h.mean(0), atol=0.04)
    assert abs(b_g.mean() - b_h.mean()) < 0.04
    np.testing.assert_allclose(th_g.std(0), th_h.std(0), rtol=0.2)


def test_poisson_gamma_model_through_the_drop_in():
    """A model no family covers (Poisson counts with per-group Gamma rates,
    data-indexed parameters): the drop-in compiles it generically and the
    group rates are recovered."""
    src = """(G : Int, N : Int, r : Real, grp : Vec Int) => {
      param shape ~ Uniform(0.5, 6.0) ;
      param lam[g] ~ Gamma(shape, r)
          for g <- 0 until G ;
      data y[n] ~ Poisson(lam[grp[n]])
          for n <- 0 until N ;
    }"""
    rng = np.random.default_rng(16)
    G, N = 6, 6000
    rates = np.array([0.5, 1.0, 2.0, 3.0, 4.5, 6.0])
    grp = rng.integers(0, G, N).astype(np.int32)
    y = rng.poisson(rates[grp]).astype(np.int32)
    with AugurInfer(None, src) as infer:
        infer.set_compile_opt(AugurOpt(target="gpu", seed=6, chains=2,
                                       initStrat=AugurUser({"shape": 2.0, "lam": np.ones(G)})))
        # the default (1.0, 0.05) overshoots: the rates' posterior sd is 0.02-0.08
        infer.set_user_sched("HMC [shape] [0.5, 0.05] (*) HMC [lam] [0.1, 0.01]")
        infer.compile(G, N, 1.0, grp)(y)
        assert infer.plan.family == "gen"
        s = infer.samplen(burnIn=200, numSamples=400)
        acc = infer.sampler.acceptance()
    lam = np.array(s["lam"]).reshape(-1, G)
    group_means = np.array([y[grp == g].mean() for g in range(G)])
    np.testing.assert_allclose(lam.mean(0), group_means, rtol=0.03)
    np.testing.assert_allclose(lam.mean(0), rates, rtol=0.12)
    assert np.all(acc > 0.3)


# ---- discrete parameters: DiscGibbs --------------------------------------------

def _mixture_pair(n, seed=21, ref_compat=True, blocks=None):
    from tests.gen_models import MIX, mixture
    hyp, dat, z = mixture(np.random.default_rng(seed), n)
    model = parse_model(MIX)
    lw = Lowering(model, hyp, dat, ref_compat, blocks=blocks)
    return lw, gm.GenModel(model, hyp, dat, ref_compat), z


def test_disc_gibbs_matches_oracle():
    """Two DiscGibbs sweeps of the mixture's assignments from a given state
    (two chains): every point's draw equals the oracle's (same Philox
    uniforms, FP64 log-space inverse CDF)."""
    print("generic DiscGibbs", flush=True)
    lw, om, z0 = _mixture_pair(3000, blocks=[("DiscGibbs", ["z"])])
    mus = [np.array([-3.5, 0.3, 4.2]), np.array([1.0, -1.0, 2.0])]
    with runtime.GenSampler(lw, [("DiscGibbs", "z")], chains=2, seed=13) as s:
        cur = []
        for c in range(2):
            s.set("mu", mus[c], c)
            s.set("z", z0, c)
            cur.append(z0.astype(np.int64))
        for sweep in range(2):
            s.step(1)
            for c in range(2):
                cur[c] = gm.disc_step(om, mus[c], cur[c], "z", 13, c, sweep, 0)
                np.testing.assert_array_equal(s.get("z", c), cur[c])
        assert len(np.unique(cur[0])) == 3


def test_mixture_sweeps_match_oracle():
    """HMC [mu] (*) DiscGibbs [z], three sweeps of two chains: the means after
    each sweep (rtol 1e-9) and every assignment equal the oracle's, and the
    log joint at the end matches it."""
    print("generic mixture sweeps", flush=True)
    blocks = [["mu"], ("DiscGibbs", ["z"])]
    lw, om, z0 = _mixture_pair(2000, blocks=blocks)
    mu0 = np.array([-3.0, 0.5, 4.0])
    with runtime.GenSampler(lw, [(["mu"], 0.3, 0.03, 0), ("DiscGibbs", "z")], chains=2, seed=17) as s:
        q = [mu0.copy(), mu0.copy() + 0.2]
        zq = [z0.astype(np.int64), z0.astype(np.int64)]
        for c in range(2):
            s.set("mu", q[c], c)
            s.set("z", z0, c)
        for sweep in range(3):
            s.step(1)
            for c in range(2):
                om.zq = zq[c]
                q[c], _ = om.hmc_step(q[c], np.ones(3, bool), 0.3, 0.03, 0, 17, c, sweep, 0, block_vars=["mu"])
                zq[c] = gm.disc_step(om, q[c], zq[c], "z", 17, c, sweep, 1)
                np.testing.assert_allclose(s.get("mu", c), q[c], rtol=1e-9)
                np.testing.assert_array_equal(s.get("z", c), zq[c])
        ll = s.log_density()
        for c in range(2):
            om.zq = zq[c]
            assert ll[c] == pytest.approx(om.ll(q[c]), rel=1e-11)


def test_mixture_through_the_drop_in():
    """A 1-D mixture with fixed weights (no family covers it): the default
    schedule (HMC on the means, DiscGibbs on the assignments) recovers the
    centres, up to label order."""
    from tests.gen_models import MIX, mixture
    centers = np.array([-4.0, 0.0, 5.0])
    hyp, dat, _ = mixture(np.random.default_rng(22), 900, centers=centers)
    with AugurInfer(None, MIX) as infer:
        infer.set_compile_opt(AugurOpt(target="gpu", seed=8, chains=2, ref_compat=False,
                                       initStrat=AugurUser({"mu": np.array([-1.0, 0.5, 1.0]),
                                                            "z": np.zeros(900, np.int32)})))
        infer.set_user_sched("HMC [mu] [0.3, 0.03] (*) DiscGibbs [z]")
        infer.compile(hyp["K"], hyp["N"], hyp["pis"], hyp["s2"])(dat["y"])
        assert infer.plan.family == "gen"
        s = infer.samplen(burnIn=300, numSamples=300)
    mu = np.sort(np.array(s["mu"]).reshape(-1, 3), axis=1)
    np.testing.assert_allclose(mu.mean(0), centers, atol=0.25)


def test_bernoulli_indicator_model():
    """z[n] ~ Bernoulli(p) used as a value (y[n] ~ Normal(3 z[n], 1)), p ~
    Beta(1, 1): the posterior mean of p is near the generating fraction."""
    from tests.gen_models import SPIKE
    rng = np.random.default_rng(23)
    N, frac = 2000, 0.3
    zt = (rng.random(N) < frac).astype(np.int32)
    y = 3.0 * zt + rng.standard_normal(N)
    with AugurInfer(None, SPIKE) as infer:
        infer.set_compile_opt(AugurOpt(target="gpu", seed=9, chains=2, ref_compat=False,
                                       initStrat=AugurUser({"p": 0.5, "z": np.zeros(N, np.int32)})))
        infer.set_user_sched("HMC [p] [0.05, 0.005] (*) DiscGibbs [z]")
        infer.compile(N)(y)
        s = infer.samplen(burnIn=200, numSamples=300)
    p = np.array(s["p"]).ravel()
    assert abs(p.mean() - zt.mean()) < 0.05
    z_last = np.array(s["z"][-1][0])
    assert (z_last == zt).mean() > 0.85
