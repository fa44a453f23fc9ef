"""The schedules of the mixture with a non-conjugate kernel for the means
(compiler/augur/test/hmvgmm2-3.infer, mvgmm2-5.infer): ESlice, HMC, MWG and
RSlice [mu], each against the oracle's sweep (the reference's kernels
restated with a per-point likelihood) and against the conjugate schedule's
posterior."""
import numpy as np
import pytest

from augurv2_amd import diagnostics

pytestmark = pytest.mark.gpu

ESLICE, HMC, MWG, RSLICE, NUTS = 1, 2, 3, 4, 5


@pytest.fixture(scope="module")
def rt():
    from augurv2_amd import _lib, runtime
    _lib.require_gpu()
    return runtime


def _data(rng, n, d, k, spread=3.0):
    pis = rng.dirichlet(np.ones(k) * 3)
    mu = rng.standard_normal((k, d)) * spread
    zt = rng.choice(k, n, p=pis)
    y = (mu[zt] + rng.standard_normal((n, d))).astype(np.float32)
    return y, zt.astype(np.int32), mu


def _settings(kernel, d):
    if kernel == HMC:
        return dict(mu_sim_len=0.5, mu_step_size=0.05)
    if kernel == RSLICE:
        return dict(mu_sim_len=0.2, mu_step_size=0.02)
    if kernel == MWG:
        return dict(mu_prop_cov=0.01 * np.eye(d))
    if kernel == NUTS:
        return dict(mu_step_size=0.01)
    return {}


@pytest.mark.parametrize("kernel,d,k,n,hier", [
    (ESLICE, 4, 3, 3000, True), (ESLICE, 4, 3, 3000, False), (ESLICE, 16, 8, 4000, True),
    (HMC, 4, 3, 3000, True), (HMC, 4, 3, 3000, False), (HMC, 16, 8, 4000, True),
    (HMC, 16, 64, 6000, False),   # 64 components x 16: every thread of the block owns a mean entry
    (MWG, 4, 3, 3000, True), (MWG, 4, 3, 3000, False),
    (RSLICE, 4, 3, 3000, True), (RSLICE, 8, 5, 3000, False),
    (NUTS, 4, 3, 3000, True), (NUTS, 4, 3, 3000, False), (NUTS, 16, 8, 2000, False),
])
def test_gmm_mu_kernel_sweeps_match_oracle(rt, oracle, kernel, d, k, n, hier):
    """Two sweeps of ConjGibbs [pis] (*) <kernel> [mu] (*) ConjGibbs [cov] (*)
    DiscGibbs [z] (or <kernel> [mu] (*) DiscGibbs [z] with pis and cov fixed)
    on the same Philox draws as the oracle.  The device evaluates the means'
    likelihood from the component statistics, the oracle point by point:
    the same function, so the moves agree to rounding."""
    rng = np.random.default_rng(100 + 10 * kernel + d + k)
    y, _, _ = _data(rng, n, d, k)
    alpha, mu0, cov0, df, scale0 = np.ones(k), np.zeros(d), 10 * np.eye(d), d + 2, np.eye(d)
    kw = _settings(kernel, d)
    C = 2
    s = rt.GmmSampler(y, alpha, mu0, cov0, df, scale0, chains=C, seed=13, sample_pi=hier,
                      sample_cov=hier, mu_kernel=kernel, ref_compat=True, **kw)
    g = oracle.Gmm(y.astype(np.float64), alpha, mu0, cov0, df, scale0, hier, hier, mu_kernel=kernel,
                   ref_compat=True, **kw)
    st = []
    for c in range(C):
        pis = rng.dirichlet(np.ones(k))
        mu = rng.standard_normal((k, d)) * 2
        cov = np.stack([np.eye(d) * (1 + 0.2 * j) for j in range(k)])
        z = rng.integers(0, k, n).astype(np.int32)
        for name, val in (("pis", pis), ("mu", mu), ("cov", cov), ("z", z)):
            s.set(name, val, chain=c)
        st.append([pis, mu, cov, z])
    for sweep in range(2):
        s.step(1)
        for c in range(C):
            pis, mu, cov, z = st[c]
            g.sweep(pis, mu, cov, z, 13, c, sweep)
            np.testing.assert_allclose(s.get("pis", c), pis, rtol=1e-9)
            np.testing.assert_allclose(s.get("mu", c), mu, rtol=1e-5, atol=1e-5)
            np.testing.assert_allclose(s.get("cov", c), cov, rtol=1e-4, atol=1e-6)
            zg = s.get("z", c)
            assert (zg == z).mean() >= 0.999
            # carry the device's state forward (the oracle's z differs in rare
            # draws within rounding of a bin edge)
            st[c] = [s.get("pis", c), s.get("mu", c), s.get("cov", c), zg.copy()]
    acc = s.mu_acceptance()
    assert np.all(acc > 0.0), acc
    s.close()


@pytest.mark.parametrize("kernel", [ESLICE, HMC, MWG, RSLICE, NUTS])
def test_gmm_mu_kernels_sample_the_conjugate_posterior(rt, kernel):
    """examples/mvgmm.py's mixture (K = 3, D = 2, fixed pis and covs), 300
    points: the posterior means of mu under each kernel equal those of
    ConjGibbs [mu] within 4 combined MCSE (at most one of 6 beyond 3)."""
    rng = np.random.default_rng(7)
    K, D, N = 3, 2, 300
    centers = np.array([[0.0, 2.0], [2.0, -2.0], [-2.0, -2.0]])
    covs = 0.1 * np.eye(D)
    y = np.stack([rng.multivariate_normal(centers[i % K], covs) for i in range(N)]).astype(np.float32)
    z0 = (np.arange(N) % K).astype(np.int32)
    C, sweeps = 4, 3000
    out = {}
    for kern in (0, kernel):
        kw = {0: {}, ESLICE: {}, HMC: dict(mu_sim_len=0.1, mu_step_size=0.01),
              MWG: dict(mu_prop_cov=0.002 * np.eye(D)), RSLICE: dict(mu_sim_len=0.05, mu_step_size=0.01),
              NUTS: dict(mu_step_size=0.01)}[kern]
        s = rt.GmmSampler(y, np.ones(K), np.zeros(D), np.eye(D), D + 2, np.eye(D), chains=C, seed=21 + kern,
                          sample_pi=False, sample_cov=False, mu_kernel=kern, ref_compat=False, **kw)
        s.set("pis", np.full(K, 1.0 / K))
        s.set("cov", np.broadcast_to(covs, (K, D, D)))
        s.set("mu", centers)
        s.set("z", z0)
        s.step(200)
        s.trace("mu", sweeps)
        s.step(sweeps)
        out[kern] = s.read_trace("mu").reshape(sweeps, C, K * D)
        s.close()
    beyond = 0
    for j in range(K * D):
        a, b = out[0][:, :, j].T, out[kernel][:, :, j].T
        assert diagnostics.rhat(b) < 1.05
        se = np.hypot(diagnostics.mcse_mean(a), diagnostics.mcse_mean(b))
        zsc = abs(a.mean() - b.mean()) / se
        assert zsc <= 4.0, (j, a.mean(), b.mean(), se)
        beyond += zsc > 3.0
    assert beyond <= 1
