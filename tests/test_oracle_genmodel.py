"""Pins of the generic-model oracle (oracle/genmodel.py) on the CPU, and
checks of the lowering's generated source (it must compile for gfx950).

The pins are:
  - ll against scipy.stats log densities.  The reference's Normal omits
    -0.5 log(2 pi), a known constant per term.
  - the reference's gradient (dotpt / dotarg chain rule) against
    central differences of ll in the textbook mode;
  - the reference-compatible Normal forms against their relation to the
    textbook ones (cbits/augur_dist.c:376 and :391-394);
  - the HMC restatement against a conjugate posterior.
"""
import math
import os

import numpy as np
import pytest
import scipy.stats as sps

from augurv2_amd.pyaugur.lower import Lowering, LoweringError
from augurv2_amd.pyaugur.rvparse import parse_model
from oracle import genmodel as gm
from tests.gen_models import models, point


def gen_model(name, ref_compat, rng=None):
    rng = rng or np.random.default_rng(7)
    src, hyp, dat = models(rng)[name]
    return gm.GenModel(parse_model(src), hyp, dat, ref_compat=ref_compat)


def test_ll_matches_scipy():
    rng = np.random.default_rng(1)
    for name in ("linreg", "hlr", "pois", "beta", "mixed"):
        m = gen_model(name, False)
        q = point(m, rng)
        env = m.unpack(q)
        ref = 0.0
        h, e = m.env, env
        if name == "linreg":
            ref += sps.invgamma.logpdf(e["s2"], h["a"], scale=h["b"])
            ref += sps.norm.logpdf(e["w"], 0.0, math.sqrt(e["s2"])).sum()
            ref += sps.norm.logpdf(e["y"], h["x"] @ e["w"], math.sqrt(e["s2"])).sum()
            n_normal = len(e["w"]) + len(e["y"])
        elif name == "hlr":
            ref += sps.expon.logpdf(e["v"], scale=1.0 / h["lam"])
            ref += sps.norm.logpdf(e["b"], 0.0, math.sqrt(e["v"]))
            ref += sps.norm.logpdf(e["theta"], 0.0, math.sqrt(e["v"])).sum()
            pr = 1 / (1 + np.exp(-(h["x"] @ e["theta"] + e["b"])))
            ref += sps.bernoulli.logpmf(e["y"], pr).sum()
            n_normal = 1 + len(e["theta"])
        elif name == "pois":
            ref += sps.uniform.logpdf(e["shape"], 0.5, 3.5)
            ref += sps.gamma.logpdf(e["lam"], e["shape"], scale=1.0 / h["r"]).sum()
            ref += sps.poisson.logpmf(e["y"], e["lam"][h["grp"]]).sum()
            n_normal = 0
        elif name == "beta":
            ref += sps.beta.logpdf(e["p"], h["a"], h["b"]) + sps.expon.logpdf(e["t"], scale=0.5)
            ref += sps.bernoulli.logpmf(e["y"], e["p"]).sum()
            ref += sps.expon.logpdf(e["u"], scale=1.0 / (2.0 * e["t"])).sum()
            n_normal = 0
        else:
            ref += sps.invgamma.logpdf(e["tau"], 3.0, scale=2.0)
            ref += sps.norm.logpdf(e["w"], 0.0, math.sqrt(e["tau"])).sum()
            mean = np.sum(e["w"][h["grp"]] * h["x"], axis=1) / (1.0 + e["tau"])
            ref += sps.norm.logpdf(e["y"], mean, math.sqrt(0.5 * e["tau"])).sum()
            n_normal = e["w"].size + len(e["y"])
        # the reference's normal_ll has no -0.5 log(2 pi)
        ref += 0.5 * math.log(2 * math.pi) * n_normal
        assert m.ll(q) == pytest.approx(ref, rel=1e-12, abs=1e-9), name


@pytest.mark.parametrize("name", ["linreg", "hlr", "pois", "beta", "mixed"])
def test_gradient_is_the_derivative_of_ll(name):
    """Textbook mode: the dotpt / dotarg chain rule equals the derivative of
    ll itself by central differences (a wrong sign, a dropped term or a
    transposed index in any dotarg fails this)."""
    rng = np.random.default_rng(2)
    m = gen_model(name, False)
    for _ in range(2):
        q = point(m, rng)
        g = m.grad(q)
        cs = np.zeros(m.P)
        for j in range(m.P):
            h = 1e-6 * max(1.0, abs(q[j]))
            qp, qm = q.copy(), q.copy()
            qp[j] += h
            qm[j] -= h
            cs[j] = (m.ll(qp) - m.ll(qm)) / (2 * h)
        np.testing.assert_allclose(g, cs, rtol=2e-6, atol=2e-6 * np.abs(cs).max())


def test_reference_normal_forms():
    """cbits/augur_dist.c:376 keeps -log var and :391-394 -0.5 var: against
    the textbook mode, each Normal term shifts ll by -0.5 log var and the
    variance gradient by (-0.5 var + 0.5 / var)."""
    rng = np.random.default_rng(3)
    ref, txt = gen_model("linreg", True), gen_model("linreg", False)
    q = point(ref, rng)
    s2 = q[0]
    n_terms = ref.layout["w"][1][0] + len(ref.env["y"])
    assert ref.ll(q) - txt.ll(q) == pytest.approx(-0.5 * math.log(s2) * n_terms, rel=1e-10)
    dg = ref.grad(q) - txt.grad(q)
    assert dg[0] == pytest.approx(n_terms * (-0.5 * s2 + 0.5 / s2), rel=1e-10)
    np.testing.assert_allclose(dg[1:], 0.0, atol=1e-12)


def test_hmc_restatement_targets_the_conjugate_posterior():
    """mu ~ Normal(m0, v0), y[n] ~ Normal(mu, s2): the posterior is
    Normal((m0/v0 + sum y/s2) V, V) with V = 1 / (1/v0 + N/s2)."""
    src = """(N : Int, m0 : Real, v0 : Real, s2 : Real) => {
      param mu ~ Normal(m0, v0) ;
      data y[n] ~ Normal(mu, s2) for n <- 0 until N ;
    }"""
    rng = np.random.default_rng(4)
    y = rng.normal(1.5, 1.0, 40)
    m = gm.GenModel(parse_model(src), dict(N=40, m0=0.0, v0=4.0, s2=1.0), dict(y=y), ref_compat=False)
    V = 1.0 / (1.0 / 4.0 + 40 / 1.0)
    mean = (y.sum() / 1.0) * V
    q = np.array([0.0])
    draws = []
    for sweep in range(1500):
        q, _ = m.hmc_step(q, [True], 0.3, 0.05, 0, seed=9, chain=0, sweep=sweep, block=0)
        if sweep >= 100:
            draws.append(q[0])
    d = np.array(draws)
    assert abs(d.mean() - mean) < 4 * math.sqrt(V / 300)
    assert 0.75 < d.var() / V < 1.3


def test_jitter_rounds_half_away_from_zero():
    # C round(): L = round(U(0.9, 1.1) simLen / eps') never rounds 2.5 down
    for sweep in range(50):
        st, L = gm.jitter(3, 0, sweep, 17, 1.0, 0.05, 0)
        assert 0.04 <= st <= 0.06 and L == max(1, int(math.floor(L + 0.5)))
        st2, L2 = gm.jitter(3, 0, sweep, 17, 1.0, 0.05, -7)
        assert st2 == st and L2 == 7
    assert gm.jitter(3, 0, 0, 17, 1.0, 0.05, 12) == (0.05, 12)


# ---- the lowering (source generation) --------------------------------------

def test_lowering_layout_and_errors():
    rng = np.random.default_rng(5)
    src, hyp, dat = models(rng)["mixed"]
    lw = Lowering(parse_model(src), hyp, dat)
    assert lw.P == 1 + 3 * 4 and lw.layout == {"tau": (0, ()), "w": (1, (3, 4))}
    # x is stored column-major: x[n][k] at k * N + n
    x = hyp["x"]
    np.testing.assert_array_equal(lw.reals[:x.size].reshape(4, -1), x.T)
    with pytest.raises(LoweringError):
        Lowering(parse_model("(a : Real) => { param p ~ Normal(0.0, a) ; "
                             "param z ~ Categorical(p) ; }"), dict(a=1.0), {})
    with pytest.raises(LoweringError, match="outside"):  # a data index past the parameter's end
        Lowering(parse_model("(G : Int, N : Int, grp : Vec Int) => { param lam[g] ~ Gamma(2.0, 1.0) "
                             "for g <- 0 until G ; data y[n] ~ Poisson(lam[grp[n]]) for n <- 0 until N ; }"),
                 dict(G=3, N=4, grp=np.array([0, 1, 3, 2])), dict(y=np.zeros(4, np.int32)))
    with pytest.raises(LoweringError, match="outside"):  # a grid longer than the indexed vector
        Lowering(parse_model("(K : Int, N : Int) => { param w[k] ~ Normal(0.0, 1.0) for k <- 0 until K ; "
                             "data y[n] ~ Normal(w[n], 1.0) for n <- 0 until N ; }"),
                 dict(K=3, N=4), dict(y=np.zeros(4)))
    with pytest.raises(LoweringError):  # a ragged grid
        Lowering(parse_model("(N : Vec Int) => { param p[d, n] ~ Normal(0.0, 1.0) "
                             "for d <- 0 until 2, n <- 0 until N[d] ; }"), dict(N=np.array([2, 3])), {})


HIPRTC = "/opt/rocm/lib/libhiprtc.so"


def _hiprtc_compile(source):
    """Compile with hipRTC exactly as augur_gen_create does (no GPU needed)."""
    import ctypes as C
    rtc = C.CDLL(HIPRTC)
    prog = C.c_void_p()
    assert rtc.hiprtcCreateProgram(C.byref(prog), source.encode(), b"pass.hip", 0, None, None) == 0
    opts = [b"--offload-arch=gfx950", b"-O3", b"-std=c++17", b"-Xclang", b"-target-feature",
            b"-Xclang", b"-packed-fp32-ops", b"-munsafe-fp-atomics"]
    arr = (C.c_char_p * len(opts))(*opts)
    r = rtc.hiprtcCompileProgram(prog, len(opts), arr)
    n = C.c_size_t(0)
    rtc.hiprtcGetProgramLogSize(prog, C.byref(n))
    log = C.create_string_buffer(n.value + 1)
    rtc.hiprtcGetProgramLog(prog, log)
    size = C.c_size_t(0)
    rtc.hiprtcGetCodeSize(prog, C.byref(size))
    rtc.hiprtcDestroyProgram(C.byref(prog))
    return r, log.value.decode(errors="replace"), size.value


@pytest.mark.skipif(not os.path.exists(HIPRTC), reason="hipRTC is not installed")
@pytest.mark.parametrize("name", ["linreg", "hlr", "pois", "beta", "mixed"])
def test_generated_source_compiles_for_gfx950(name):
    rng = np.random.default_rng(6)
    src, hyp, dat = models(rng)[name]
    for ref_compat in (True, False):
        lw = Lowering(parse_model(src), hyp, dat, ref_compat)
        r, log, size = _hiprtc_compile(lw.source)
        assert r == 0 and size > 0, log[-2000:]


# ---- discrete parameters (DiscGibbs) ------------------------------------------

def test_disc_step_draws_the_exact_conditional():
    """One point of the mixture, fixed means: p(z = k | y) is proportional
    to pis_k N(y | mu_k, s2).  The oracle's DiscGibbs draws over 4000 sweeps
    (fresh Philox counters) match those probabilities within 4 standard
    errors; a wrong normalisation, a shifted bin or a dropped term fails."""
    from tests.gen_models import MIX
    hyp = dict(K=3, N=1, pis=np.array([0.2, 0.5, 0.3]), s2=2.0)
    y = np.array([0.7])
    m = gm.GenModel(parse_model(MIX), hyp, dict(y=y), ref_compat=True)
    mu = np.array([-1.0, 0.5, 2.0])
    w = hyp["pis"] * np.exp(-0.5 * (y[0] - mu) ** 2 / hyp["s2"])
    prob = w / w.sum()
    counts = np.zeros(3)
    zq = np.zeros(1, np.int64)
    for sweep in range(4000):
        zq = gm.disc_step(m, mu, zq, "z", seed=5, chain=0, sweep=sweep, block=1)
        counts[zq[0]] += 1
    freq = counts / counts.sum()
    se = np.sqrt(prob * (1 - prob) / 4000)
    assert np.all(np.abs(freq - prob) < 4 * se), (freq, prob)


def test_disc_lowering_errors():
    from tests.gen_models import MIX
    hyp = dict(K=3, N=4, pis=np.full(3, 1 / 3), s2=1.0)
    dat = dict(y=np.zeros(4))
    # a term using z at another point cannot be drawn point by point
    src = MIX.replace("Normal(mu[z[n]], s2)", "Normal(mu[z[n]] + mu[z[0]], s2)")
    with pytest.raises(LoweringError, match="another point"):
        Lowering(parse_model(src), hyp, dat, blocks=[["mu"], ("DiscGibbs", ["z"])])
    # the pmf of a discrete parameter is a fixed vector
    src = MIX.replace("Categorical(pis)", "Categorical(mu)")
    with pytest.raises(LoweringError, match="pmf"):
        Lowering(parse_model(src), hyp, dat)
    with pytest.raises(LoweringError, match="DiscGibbs updates"):
        Lowering(parse_model(MIX), hyp, dat, blocks=[("DiscGibbs", ["mu"])])


@pytest.mark.skipif(not os.path.exists(HIPRTC), reason="hipRTC is not installed")
def test_discrete_models_compile_for_gfx950():
    from tests.gen_models import MIX, SPIKE, mixture
    hyp, dat, _ = mixture(np.random.default_rng(3), 200)
    for src, h, d, blocks in ((MIX, hyp, dat, [["mu"], ("DiscGibbs", ["z"])]),
                              (SPIKE, dict(N=50), dict(y=np.zeros(50)), [["p"], ("DiscGibbs", ["z"])])):
        lw = Lowering(parse_model(src), h, d, blocks=blocks)
        assert [k[2] for k in lw.block_kernels] == ["HMC", "DiscGibbs"]
        r, log, size = _hiprtc_compile(lw.source)
        assert r == 0 and size > 0, log[-2000:]
