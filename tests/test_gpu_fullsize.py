"""Parity at the full sizes of the benchmark configurations, where the FP64
oracle still finishes in seconds: the LDA corpus (K = 10, V = 1000, D = 200,
~20k tokens) sweep by sweep, the GMM model log density at N = 1M, D = 16,
K = 32 after a sweep, and the HLR data term (log density and gradient in
FP32, log density in FP64 on the matrix cores) at N = 1M, D = 128, 8
chains."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    from augurv2_amd import _lib, runtime
    _lib.require_gpu()
    return runtime


def test_lda_benchmark_corpus_matches_oracle(rt, oracle):
    """The LDA benchmark corpus, 3 sweeps: theta and phi to FP64 rounding,
    count tables through them, z agreeing except within rounding of a bin
    edge (each sweep continues from the GPU's assignment)."""
    from augurv2_amd import workloads
    doc_off, w, alpha, beta, _ = workloads.lda_corpus()
    K = len(alpha)
    rng = np.random.default_rng(4)
    z = rng.integers(0, K, len(w)).astype(np.int32)
    s = rt.LdaSampler(K, doc_off, w, alpha, beta, chains=1, seed=12)
    s.set("z", z)
    D, V = len(doc_off) - 1, len(beta)
    for sweep in range(3):
        s.step(1)
        theta, phi, zr = np.zeros((D, K)), np.zeros((K, V)), z.copy()
        oracle.lda_sweep(doc_off, w, alpha, beta, 12, 0, sweep, theta, phi, zr)
        np.testing.assert_allclose(s.get("theta"), theta, rtol=1e-9, atol=1e-300)
        np.testing.assert_allclose(s.get("phi"), phi, rtol=1e-9, atol=1e-300)
        zg = s.get("z")
        assert (zg == zr).mean() >= 0.999
        z = zg
    s.close()


def test_gmm_benchmark_size_log_density(rt, oracle):
    """N = 1M, D = 16, K = 32, 2 chains, one sweep from the generating
    parameters: the model log density (prior and data terms) equals the
    oracle's at the GPU's state within 2e-6 relative (FP32 data terms);
    assignments stay in range and the weights sum to one."""
    from augurv2_amd import workloads
    y, zt, pis_t, mu_t, cov_t = workloads.gmm_data()
    n, d = y.shape
    k, C, df = len(pis_t), 2, d + 2
    alpha, mu0, cov0, scale0 = np.ones(k), np.zeros(d), 10.0 * np.eye(d), np.eye(d)
    s = rt.GmmSampler(y, alpha, mu0, cov0, df, scale0, chains=C, seed=3)
    for c in range(C):
        s.set("pis", pis_t, chain=c)
        s.set("mu", mu_t, chain=c)
        s.set("cov", cov_t, chain=c)
        s.set("z", zt.astype(np.int32), chain=c)
    s.step(1)
    ll = s.log_density()
    g = oracle.Gmm(y.astype(np.float64), alpha, mu0, cov0, df, scale0)
    for c in range(C):
        pis, mu, cov, z = s.get("pis", c), s.get("mu", c), s.get("cov", c), s.get("z", c)
        assert z.min() >= 0 and z.max() < k
        assert abs(pis.sum() - 1.0) < 1e-12
        ref = g.ll_mod(pis, mu, cov, z)
        assert abs(ll[c] - ref) <= 2e-6 * abs(ref), (c, ll[c], ref)
    s.close()


def test_hlr_data_term_at_surrogate_size(oracle):
    """N = 1M rows, D = 128, 8 chains: the FP32 pass (log density within 1e-5
    relative, gradient within 1e-5 of sum |a x| per component) and the FP64
    matrix-core log density (1e-11 relative) against the FP64 oracle."""
    from augurv2_amd import _lib, ops
    _lib.require_gpu()
    rng = np.random.default_rng(9)
    n, d, C = 1_000_000, 128, 8
    x = rng.standard_normal((n, d), dtype=np.float32)
    beta_true = rng.standard_normal(d) * 0.1
    y = (rng.random(n) < 1 / (1 + np.exp(-(x.astype(np.float64) @ beta_true)))).astype(np.int32)
    beta = np.concatenate([rng.standard_normal((C, d)) * 0.1, np.zeros((C, 1))], axis=1)
    ll, g = ops.hlr_ll_grad(x, y, beta)
    ll64 = ops.hlr_ll64(x, y, beta)
    x64 = x.astype(np.float64)
    m = oracle.Hlr(x64, y)
    for c in range(C):
        st = np.concatenate([beta[c], [1.0]])
        ref = m.data_ll(st)
        assert abs(ll[c] - ref) <= 1e-5 * abs(ref)
        assert abs(ll64[c] - ref) <= 1e-11 * abs(ref)
        g_ref, _ = m.data_grad(st)
        t = x64 @ beta[c, :d]
        a = np.abs(y - 1 / (1 + np.exp(-t)))
        scale = np.abs(x64).T @ a
        assert np.all(np.abs(g[c, :d] - g_ref) <= 1e-5 * scale + 1e-9)
