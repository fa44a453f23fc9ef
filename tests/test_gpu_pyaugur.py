"""The pyaugur drop-in end to end on the GPU, driven exactly like the
reference's examples (examples/lda.py, examples/hlr.py, examples/mvgmm.py)."""
import numpy as np
import pytest

from augurv2_amd.pyaugur.augurlib import AugurDefault, AugurInfer, AugurOpt, AugurUser

pytestmark = pytest.mark.gpu

LDA = """(K : Int, D : Int, N : Vec Int, alpha : Vec Real, beta : Vec Real) => {
  param theta[d] ~ Dirichlet(alpha)
      for d <- 0 until D ;
  param phi[k] ~ Dirichlet(beta)
      for k <- 0 until K ;
  param z[d, n] ~ Categorical(theta[d])
      for d <- 0 until D, n <- 0 until N[d] ;
  data w[d, n] ~ Categorical(phi[z[d, n]])
      for d <- 0 until D, n <- 0 until N[d] ;
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

MVGMM = """
(K : Int, N : Int, mu0 : Vec Real, covs0 : Mat Real, pis : Vec Real, covs : Mat Real) => {
  param mu[k] ~ MvNormal(mu0, covs0)
      for k <- 0 until K ;
  param z[n] ~ Categorical(pis)
      for n <- 0 until N ;
  data y[n] ~ MvNormal(mu[z[n]], covs)
      for n <- 0 until N ;
}
"""


@pytest.fixture(scope="module", autouse=True)
def gpu():
    from augurv2_amd import _lib
    _lib.require_gpu()


def test_lda_example():
    # examples/lda.py:41-76 (D = 4, K = 4, V = 8, designed topics)
    K, D, V = 4, 4, 8
    N = np.array([10, 10, 10, 10], dtype=np.int32)
    alpha, beta = np.full(K, 0.1), np.full(V, 0.01)
    w = np.array([[0, 2, 1, 3, 0, 2, 1, 3, 0, 2], [4, 6, 5, 7, 4, 6, 5, 7, 5, 7],
                  [2, 4, 3, 5, 2, 4, 3, 5, 2, 4], [0, 6, 1, 7, 0, 6, 1, 7, 1, 7]], dtype=np.int32)
    with AugurInfer("config.yml", LDA) as infer:
        infer.set_compile_opt(AugurOpt(cached=False, target="gpu", paramScale=None, seed=1))
        infer.set_user_sched("ConjGibbs [theta] (*) ConjGibbs [phi] (*) DiscGibbs [z]")
        infer.compile(K, D, N, alpha, beta)(w)
        samples = infer.samplen(burnIn=200, numSamples=300)
        assert set(samples) == {"theta", "phi", "z"}
        assert len(samples["theta"]) == 300
        th = samples["theta"][-1]
        assert th.dtype == object and len(th) == D and np.allclose(th[0].sum(), 1.0)
        assert len(samples["z"][-1]) == D and samples["z"][-1][0].dtype == np.int32
        # words 2 and 3 (one designed topic) share their most likely topic
        top = np.array([np.argmax(np.stack(phi), axis=0) for phi in samples["phi"]])
        assert np.mean(top[:, 2] == top[:, 3]) > 0.9
        ll = infer.curr_ll()
        assert np.isfinite(ll)


def test_hlr_example_schedule():
    rng = np.random.default_rng(0)
    K, N, lam = 5, 3000, 1.0
    x = rng.standard_normal((N, K))
    theta = np.array([1.0, -1.0, 0.5, 0.0, 2.0])
    y = (rng.random(N) < 1 / (1 + np.exp(-(x @ theta - 0.5)))).astype(np.int32)
    with AugurInfer("config.yml", HLR) as infer:
        opt = AugurOpt(target="gpu", initStrat=AugurUser({"v": 1.0, "b": 0.0,
                                                          "theta": np.zeros(K)}), seed=2)
        infer.set_compile_opt(opt)
        infer.set_user_sched("HMC [v] [0.5, 0.05] (*) HMC [b] [0.5, 0.05] (*) HMC [theta] [0.5, 0.05]")
        infer.compile(K, N, lam, x)(y)
        s = infer.samplen(burnIn=300, numSamples=500)
        assert isinstance(s["v"][0], float) and s["theta"][0].shape == (K,)
        post = np.mean(s["theta"], axis=0)
        assert np.all(np.abs(post - theta) < 0.25), post
        assert abs(np.mean(s["b"]) + 0.5) < 0.25


DETHLR = """(K : Int, N : Int, lam : Real, x : Vec (Vec Real)) => {
  param v ~ Exponential(lam) ;
  param b ~ Normal(0.0, v) ;
  param theta[k] ~ Normal(0.0, v)
      for k <- 0 until K ;
  param bias[n] : Real = sigmoid(dotprod(theta, x[n]) + b)
      for n <- 0 until N ;
  data y[n] ~ Bernoulli(bias[n])
      for n <- 0 until N ;
}
"""


@pytest.mark.parametrize("sched", [None, "MWG [v] ~ Normal(v, 1.0) (*) MWG [b] ~ Normal(b, 0.01) (*) HMC [theta]"])
def test_dethlr_deterministic_bias(sched):
    """compiler/augur/test/dethlr.rv (the Bernoulli bias a deterministic
    parameter) with the automatic schedule and with hlr2.infer's MWG kernels:
    the samples hold bias[n] = sigmoid(x[n] . theta + b) at every kept sweep,
    and b is recovered."""
    rng = np.random.default_rng(5)
    K, N = 6, 3000
    x = rng.standard_normal((N, K))
    theta, b = rng.standard_normal(K) * 0.5, -0.5
    y = (rng.random(N) < 1 / (1 + np.exp(-(x @ theta + b)))).astype(np.int32)
    xs = np.empty(N, dtype=object)
    for i in range(N):
        xs[i] = x[i]
    with AugurInfer("config.yml", DETHLR) as infer:
        infer.set_compile_opt(AugurOpt(target="gpu", seed=3, chains=1, ref_compat=False,
                                       initStrat=AugurUser({"v": 1.0, "b": 0.0, "theta": np.zeros(K)})))
        if sched:
            infer.set_user_sched(sched)
        infer.compile(K, N, 1.0, xs)(y)
        s = infer.samplen(burnIn=300, numSamples=50)
        assert set(s) == {"v", "b", "theta", "bias"}
        for i in (0, 49):
            ref = 1 / (1 + np.exp(-(x.astype(np.float32).astype(np.float64) @ s["theta"][i] + s["b"][i])))
            np.testing.assert_allclose(s["bias"][i], ref, rtol=1e-12)
        assert abs(np.mean(s["b"]) - b) < 0.25


@pytest.mark.parametrize("sched", [
    "ConjGibbs [mu] (*) DiscGibbs [z]",                                  # examples/mvgmm.py sched1
    "ESlice [mu] (*) DiscGibbs [z]",                                     # sched2
    "HMC [mu] [0.2, 0.01] (*) DiscGibbs [z]",                            # sched3
    "MWG [mu] ~ MvNormal(mu[k], covs0) for k <- 0 until K (*) DiscGibbs [z]",  # sched4
    "RSlice [mu] [0.05, 0.01] (*) DiscGibbs [z]",                        # mvgmm5.infer
    "NUTS [mu] [0.01] (*) DiscGibbs [z]",                                # mvgmm6.infer
])
def test_mvgmm_example(sched):
    # examples/mvgmm.py:39-73 (K = 3, D = 2, fixed pis and covs) with each of
    # its schedules
    rng = np.random.default_rng(1)
    K, D, N = 3, 2, 300
    mu0, covs0, pis = np.zeros(D), np.eye(D), np.full(K, 1.0 / K)
    covs = np.array([[0.1, 0.0], [0.0, 0.1]])
    centers = [np.array([0.0, 2.0]), np.array([2.0, -2.0]), np.array([-2.0, -2.0])]
    y = np.array([rng.multivariate_normal(centers[n % K], covs) for n in range(N)])
    opt = AugurOpt(target="gpu", seed=4, chains=4)
    if sched.startswith("MWG"):
        # the proposal MvNormal(mu[k], covs0) has 30x the posterior's spread
        # and accepts about one move in a thousand: it holds the centres (from
        # which it starts here) rather than finding them
        opt = AugurOpt(target="gpu", seed=4, chains=4, initStrat=AugurUser(
            {"mu": np.array(centers), "z": (np.arange(N) % K).astype(np.int32)}))
    with AugurInfer("config.yml", MVGMM) as infer:
        infer.set_compile_opt(opt)
        infer.set_user_sched(sched)
        infer.compile(K, N, mu0, covs0, pis, covs)(y)
        s = infer.samplen(burnIn=150, numSamples=20)
        ll = infer.curr_ll()
        # Gibbs from a random start can sit in a local mode (two components on
        # one cluster); the chain of highest log density holds the global one,
        # which puts one centre on each cluster (up to label order)
        best = int(np.argmax(ll))
        mus = np.stack(s["mu"][-1][best])
        found = sorted(map(tuple, np.round(mus)))
        assert found == sorted(map(tuple, np.round(np.array(centers)))), (found, ll)


def test_multichain_and_set_pt():
    K, D, V = 3, 2, 5
    N = np.array([4, 3], np.int32)
    w = np.array([np.array([0, 1, 2, 3], np.int32), np.array([4, 0, 1], np.int32)], dtype=object)
    with AugurInfer(None, LDA) as infer:
        infer.set_compile_opt(AugurOpt(target="gpu", chains=3, initStrat=AugurDefault()))
        infer.compile(K, D, N, np.ones(K), np.ones(V))(w)
        s = infer.samplen(numSamples=2)
        assert len(s["z"][0]) == 3  # per-chain values
        infer.set_pt({"theta": [np.full(K, 1 / K)] * D, "phi": [np.full(V, 1 / V)] * K,
                      "z": [np.zeros(4, np.int32), np.ones(3, np.int32)]})
        one = infer.sample()
        assert len(one["theta"]) == 3
