"""The experiments harness on the device: predictive log-likelihood terms
against the reference's loops (oracle/predll.py), the sampler-state forms,
and a run that writes the reference's stats.yml."""
import numpy as np
import pytest

from augurv2_amd import experiments
from oracle import predll

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    from augurv2_amd import _lib, runtime
    _lib.require_gpu()
    return runtime


def test_lda_predictive_term_matches_the_reference_loop(rt):
    rng = np.random.default_rng(60)
    D, K, V = 9, 6, 40
    theta = rng.dirichlet(np.ones(K) * 0.3, D)
    phi = rng.dirichlet(np.ones(V) * 0.1, K)
    w_ho = [rng.integers(0, V, n) for n in (3, 0, 7, 1, 12, 0, 5, 9, 2)]
    got = experiments.lda_predictive_ll_term(theta, phi, w_ho)
    ref = predll.lda_term(theta, phi, w_ho)
    assert abs(got - ref) <= 1e-12 * abs(ref)


@pytest.mark.parametrize("k,d,n", [(4, 3, 500), (32, 16, 300)])
def test_gmm_predictive_term_matches_the_reference_loop(rt, k, d, n):
    from scipy.stats import invwishart
    rng = np.random.default_rng(61 + k)
    pis = rng.dirichlet(np.ones(k))
    mu = rng.standard_normal((k, d)) * 3
    cov = np.stack([invwishart.rvs(d + 4, np.eye(d), random_state=rng) for _ in range(k)])
    y = rng.standard_normal((n, d)) * 3
    got = experiments.gmm_predictive_ll_term(pis, mu, cov, y)
    ref = predll.gmm_term(pis, mu, cov, y)
    assert abs(got - ref) <= 1e-10 * abs(ref)


def test_sampler_state_forms(rt):
    """augur_lda_predictive_ll / augur_gmm_predictive_ll at each chain's
    current state equal the term at the parameters read back."""
    rng = np.random.default_rng(62)
    K, D, V, C = 5, 12, 60, 2
    lens = rng.integers(5, 30, D)
    doc_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    w = rng.integers(0, V, doc_off[-1]).astype(np.int32)
    s = rt.LdaSampler(K, doc_off, w, np.full(K, 0.2), np.full(V, 0.05), chains=C, seed=3)
    s.set("z", rng.integers(0, K, len(w)).astype(np.int32))
    s.step(3)
    ho = [rng.integers(0, V, n) for n in rng.integers(0, 6, D)]
    off, flat = experiments._csr(ho)
    got = s.predictive_ll(off, flat)
    for c in range(C):
        ref = predll.lda_term(s.get("theta", c), s.get("phi", c), ho)
        assert abs(got[c] - ref) <= 1e-12 * abs(ref)
    s.close()
    n, d, k = 2000, 4, 3
    mu = rng.standard_normal((k, d)) * 3
    y = (mu[rng.integers(0, k, n)] + rng.standard_normal((n, d))).astype(np.float32)
    g = rt.GmmSampler(y, np.ones(k), np.zeros(d), 10 * np.eye(d), d + 2, np.eye(d), chains=C, seed=4)
    g.set("mu", mu)
    g.set("z", rng.integers(0, k, n).astype(np.int32))
    g.step(4)
    yt = mu[rng.integers(0, k, 200)] + rng.standard_normal((200, d))
    got = g.predictive_ll(yt)
    for c in range(C):
        ref = predll.gmm_term(g.get("pis", c), g.get("mu", c), g.get("cov", c), yt)
        assert abs(got[c] - ref) <= 1e-10 * abs(ref)
    g.close()


def test_hmvgmm_experiment_writes_stats(rt, tmp_path):
    """An hmvgmm run as experiments/hmvgmm/hmvgmm.py drives it: compile and
    sample timed, the held-out log-likelihood's running mean from the
    initial point over the samples, stats.yml written; the predictive
    log-likelihood rises from the initial point."""
    from augurv2_amd.pyaugur.augurlib import AugurInfer, AugurOpt
    model = open_model()
    rng = np.random.default_rng(63)
    K, D, N = 3, 2, 600
    centers = np.array([[0.0, 4.0], [4.0, -4.0], [-4.0, -4.0]])
    y = centers[np.arange(N) % K] + rng.standard_normal((N, D)) * 0.5
    y_test = centers[np.arange(150) % K] + rng.standard_normal((150, D)) * 0.5
    hyp = (K, N, np.ones(K), np.zeros(D), 10 * np.eye(D), D + 2, np.eye(D))
    timer = experiments.Timed()
    with AugurInfer(None, model) as infer:
        infer.set_compile_opt(AugurOpt(target="gpu", seed=2))
        timer.compile(infer, hyp, (y,))
        init = infer.sample()
        S = 20
        out = timer.samplen(infer, burnIn=100, numSamples=S)
    pred = experiments.gmm_predictive_ll(out["pis"], out["mu"], out["cov"], y_test, init, S)
    assert np.isfinite(pred).all() and pred[-1] > pred[0]
    p = experiments.dump_pred_ll_stats(str(tmp_path / "stat"), timer.compile_time, timer.sample_time, pred)
    import yaml
    st = yaml.safe_load(open(p))
    assert st["compile"] > 0 and st["sample"] > 0 and len(st["predll"]) == S + 1


def open_model():
    return """(K : Int, N : Int, alpha : Vec Real, mu0 : Vec Real, covs0 : Mat Real, df : Int, scale0 : Mat Real) => {
  param pis ~ Dirichlet(alpha) ;
  param mu[k] ~ MvNormal(mu0, covs0)
      for k <- 0 until K ;
  param cov[k] ~ InvWishart(df, scale0)
      for k <- 0 until K ;
  param z[n] ~ Categorical(pis)
      for n <- 0 until N ;
  data y[n] ~ MvNormal(mu[z[n]], cov[z[n]])
      for n <- 0 until N ;
}
"""
