"""Kernel-level parity: each HIP kernel (through the C-ABI) against the CPU
FP64 restatement on the same seeded inputs and the same Philox streams."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ops():
    from augurv2_amd import _lib, ops as o
    _lib.require_gpu()
    return o


KAT = [  # Random123 known-answer vectors for philox4x32-10
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


def test_philox_kat_and_random(ops, oracle):
    ctr = np.array([k[0] for k in KAT], np.uint32)
    key = np.array([k[1] for k in KAT], np.uint32)
    out = ops.philox(ctr, key)
    assert (out == np.array([k[2] for k in KAT], np.uint32)).all()
    rng = np.random.default_rng(1)
    ctr = rng.integers(0, 2**32, (5000, 4), dtype=np.uint64).astype(np.uint32)
    key = rng.integers(0, 2**32, (5000, 2), dtype=np.uint64).astype(np.uint32)
    out = ops.philox(ctr, key)
    ref = np.array([oracle.philox(c, k) for c, k in zip(ctr[:300], key[:300])])
    assert (out[:300] == ref).all()


def test_gamma_matches_oracle(ops, oracle):
    shapes = np.array([0.01, 0.1, 0.5, 0.999, 1.0, 1.5, 2.5, 10.0, 123.4, 1e4, 1e6] * 50)
    out = ops.gamma(7, 3, 11, 2, shapes)
    ref = np.array([oracle.gamma(7, 3, i, 11, 2, s) for i, s in enumerate(shapes)])
    assert np.isfinite(out).all()
    # FP64 on both sides; only libm ulp differences remain
    np.testing.assert_allclose(out, ref, rtol=1e-11, atol=0)


def test_gamma_rejects_bad_shape(ops):
    out = ops.gamma(0, 0, 0, 1, np.array([0.0, -1.0, np.nan, np.inf]))
    assert np.isnan(out).all()


def test_dirichlet_conj_matches_oracle(ops, oracle):
    rng = np.random.default_rng(2)
    k = 10
    alpha = np.full(k, 0.1)
    cnt = rng.integers(0, 50, (37, k)).astype(np.int32)
    cnt[3] = 0  # an empty row
    out = ops.dirichlet_conj(5, 1, 9, oracle.TAG_LDA_THETA, alpha, cnt)
    ref = np.array([oracle.dirichlet_conj(5, 1, 9, oracle.TAG_LDA_THETA, r * k, alpha, cnt[r])
                    for r in range(cnt.shape[0])])
    np.testing.assert_allclose(out, ref, rtol=1e-10, atol=1e-300)
    np.testing.assert_allclose(out.sum(1), 1.0, rtol=1e-12)


def _corpus(rng, D, K, V, mean_len):
    lens = np.maximum(rng.poisson(mean_len, D), 0)
    lens[0] = 0  # an empty document
    doc_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    w = rng.integers(0, V, doc_off[-1]).astype(np.int32)
    return doc_off, w


@pytest.mark.parametrize("K", [4, 10, 64, 100, 120])
def test_lda_zstep_bit_exact(ops, oracle, K):
    """z bit-exact with the oracle's FP32 draw and both count tables exact.
    Up to K * V = 32768 the kernel counts word-topic pairs in an LDS table
    per block (K <= 100 here); K = 120 (36,000 cells) takes the form that
    adds every token in HBM; K = 100 also takes the per-token topic atomics
    (K > 64)."""
    rng = np.random.default_rng(K)
    D, V = 57, 300
    doc_off, w = _corpus(rng, D, K, V, 150)
    theta = rng.dirichlet(np.full(K, 0.1), D).astype(np.float32)
    phi = rng.dirichlet(np.full(V, 0.01), K)
    phi_t = np.ascontiguousarray(phi.T).astype(np.float32)
    z, ct, cp = ops.lda_zstep(doc_off, w, theta, phi_t, V, 42, 2, 17)
    zref = oracle.lda_zstep_f32(doc_off, w, theta, phi_t, 42, 2, 17)
    assert (z == zref).all()
    ct_ref, cp_ref = oracle.lda_counts(D, K, V, doc_off, w, z)
    assert (ct == ct_ref).all() and (cp == cp_ref).all()
    assert ct.sum() == len(w) and cp.sum() == len(w)


def test_lda_zstep_vs_fp64_reference_form(ops, oracle):
    """The FP32 kernel draw and the reference FP64 draw (augur_unnorm_disc_samp)
    with the same uniform differ only where the uniform sits on a boundary."""
    rng = np.random.default_rng(5)
    D, K, V = 80, 10, 500
    doc_off, w = _corpus(rng, D, K, V, 200)
    theta = rng.dirichlet(np.full(K, 0.5), D).astype(np.float32)
    phi_t = np.ascontiguousarray(rng.dirichlet(np.full(V, 0.1), K).T).astype(np.float32)
    z, _, _ = ops.lda_zstep(doc_off, w, theta, phi_t, V, 9, 0, 3)
    lib = oracle.lib()
    mism = 0
    for d in range(D):
        for t in range(doc_off[d], doc_off[d + 1]):
            pmf = theta[d].astype(np.float64) * phi_t[w[t]].astype(np.float64)
            u = float(oracle.lib().or_u01f(int(oracle.draw4(9, 0, t, 0, 3, oracle.TAG_LDA_Z)[0])))
            zr = lib.or_categorical_ref(np.ascontiguousarray(pmf), K, u)
            if zr != z[t]:
                mism += 1
                cdf = np.cumsum(pmf) / pmf.sum()
                assert np.min(np.abs(cdf - u)) < 1e-5
    assert mism <= 2


@pytest.mark.parametrize("lds", [False, True])
@pytest.mark.parametrize("n,d,chains,intercept", [(5000, 16, 4, False), (4099, 64, 4, True),
                                                  (1000, 128, 8, False), (777, 5, 3, True),
                                                  (100003, 64, 4, False), (3001, 32, 3, True),
                                                  (2500, 16, 2, False), (129, 64, 1, True),
                                                  (7, 32, 4, False), (20000, 64, 2, False),
                                                  (50001, 128, 8, False), (3000, 128, 3, True),
                                                  (130, 128, 1, False), (9000, 128, 2, True)])
def test_hlr_ll_grad_vs_fp64(ops, oracle, monkeypatch, n, d, chains, intercept, lds):
    """Register form of the pass (D <= 64 with chains <= D / 8, and D = 128
    with up to 8 chains, padded to a power of two) and the LDS-staged form
    forced on the same shapes; partial tiles, a block of fewer rows than
    lanes, the intercept."""
    monkeypatch.setenv("AUGUR_HLR_PASS_LDS", "1" if lds else "0")
    rng = np.random.default_rng(n + d)
    x = rng.standard_normal((n, d)).astype(np.float32)
    beta_true = rng.standard_normal(d) * 0.3
    y = (rng.random(n) < 1 / (1 + np.exp(-(x @ beta_true)))).astype(np.int32)
    beta = rng.standard_normal((chains, d + 1)) * 0.3
    if not intercept:
        beta[:, d] = 0.0
    ll, g = ops.hlr_ll_grad(x, y, beta, intercept)
    m = oracle.Hlr(x.astype(np.float64), y, has_intercept=intercept)
    x64 = x.astype(np.float64)
    for c in range(chains):
        st = np.concatenate([beta[c], [1.0]])
        ll_ref = m.data_ll(st)
        g_ref, gb_ref = m.data_grad(st)
        # tolerance 1e-5 relative (FP32 arithmetic, FP64 accumulation)
        assert abs(ll[c] - ll_ref) <= 1e-5 * abs(ll_ref)
        t = x64 @ beta[c, :d] + beta[c, d]
        a = np.abs(y - 1 / (1 + np.exp(-t)))
        scale = np.abs(x64).T @ a  # sum_n |a_n x_nj|: the conditioning of each component
        assert np.all(np.abs(g[c, :d] - g_ref) <= 1e-5 * scale + 1e-9)
        if intercept:
            assert abs(g[c, d] - gb_ref) <= 1e-5 * a.sum() + 1e-9


@pytest.mark.parametrize("chains", [16, 2])
@pytest.mark.parametrize("y", [0, 1])
def test_hlr_row_term_per_value(ops, chains, y):
    """The FP32 row term on its own: one row x = e_0, so chain c's
    log-density is log sigmoid(+-t_c) and its gradient the adjoint
    y - sigmoid(t_c) at t_c = theta_c0 (exact in FP32).  Against the stable
    FP64 forms: both within 8 ulp relative (LDS form at 16 chains, register
    form at 2), over |t| up to 100 and near 0."""
    d = 16
    x = np.zeros((1, d), np.float32)
    x[0, 0] = 1.0
    ts = np.concatenate([np.linspace(-30, 30, 61), [0.0, 1e-8, -1e-8, 1e-3, -1e-3, 0.37, -0.37, 17.0,
                                                   -17.0, 40.0, -40.0, 88.0, -88.0, 100.0, -100.0]])
    ts = ts.astype(np.float32).astype(np.float64)
    for k in range(0, len(ts), chains):
        tc = ts[k:k + chains]
        beta = np.zeros((len(tc), d + 1))
        beta[:, 0] = tc
        ll, g = ops.hlr_ll_grad(x, np.array([y], np.int32), beta)
        st = tc if y else -tc  # log sigmoid(st) = min(st, 0) - log1p(exp(-|st|))
        ll_ref = np.minimum(st, 0.0) - np.log1p(np.exp(-np.abs(st)))
        # a = y - sigmoid(t): sigmoid(-t) for y = 1, -sigmoid(t) for y = 0
        def sig(s):
            return np.where(s >= 0, 1 / (1 + np.exp(-np.abs(s))),
                            np.exp(-np.abs(s)) / (1 + np.exp(-np.abs(s))))
        a_ref = sig(-tc) if y else -sig(tc)
        # 8 ulp relative; values below FP32's normal range (|t| near 100) absolutely
        tol, tiny = 8 * 2.0**-24, 2.0**-126
        np.testing.assert_allclose(ll, ll_ref, rtol=tol, atol=tiny, err_msg=f"t {tc}")
        np.testing.assert_allclose(g[:, 0], a_ref, rtol=tol, atol=tiny, err_msg=f"t {tc}")


@pytest.mark.parametrize("n,d,chains,intercept", [(5000, 16, 4, False), (4099, 64, 4, False),
                                                  (1000, 128, 8, False), (2003, 128, 16, False),
                                                  (15, 32, 2, False), (777, 5, 3, True),
                                                  (300, 64, 4, True)])
def test_hlr_ll64_vs_fp64(ops, oracle, n, d, chains, intercept):
    """FP64 row arithmetic (matrix-core form for D in {16, 32, 64, 128}
    without intercept, including row counts that leave a partial 16-row
    tile): the oracle's FP64 log-density to 1e-11 relative."""
    rng = np.random.default_rng(7 * n + d)
    x = rng.standard_normal((n, d)).astype(np.float32)
    beta_true = rng.standard_normal(d) * 0.3
    y = (rng.random(n) < 1 / (1 + np.exp(-(x @ beta_true)))).astype(np.int32)
    beta = rng.standard_normal((chains, d + 1)) * 0.3
    if not intercept:
        beta[:, d] = 0.0
    ll = ops.hlr_ll64(x, y, beta, intercept)
    m = oracle.Hlr(x.astype(np.float64), y, has_intercept=intercept)
    for c in range(chains):
        ll_ref = m.data_ll(np.concatenate([beta[c], [1.0]]))
        assert abs(ll[c] - ll_ref) <= 1e-11 * abs(ll_ref), (c, ll[c], ll_ref)


@pytest.mark.parametrize("n,d,chains", [(4099, 16, 4), (2000, 128, 8), (1600, 64, 16)])
def test_hlr_ll64_terms_over_a_wide_range(ops, oracle, n, d, chains):
    """Row terms log sigmoid(+-t) with t spread over [-60, 60] (both branches
    of the FP64 log(1 + e^-x) evaluation, its switch at x = 0.8814, and terms
    far below one ulp of the sum): each chain's sum within 1e-13 relative of
    the oracle's libm form."""
    rng = np.random.default_rng(n + d)
    x = (rng.standard_normal((n, d)) * 0.01).astype(np.float32)
    x[:, 0] = np.linspace(-60.0, 60.0, n).astype(np.float32)
    y = (rng.random(n) < 0.5).astype(np.int32)
    beta = np.zeros((chains, d + 1))
    beta[:, 0] = (np.arange(chains) + 1.0) / chains
    beta[:, 1:d] = rng.standard_normal((chains, d - 1)) * 0.3
    ll = ops.hlr_ll64(x, y, beta, False)
    m = oracle.Hlr(x.astype(np.float64), y)
    for c in range(chains):
        ll_ref = m.data_ll(np.concatenate([beta[c], [1.0]]))
        assert abs(ll[c] - ll_ref) <= 1e-13 * abs(ll_ref), (c, ll[c], ll_ref)


def _gmm_params(rng, k, d):
    pis = rng.dirichlet(np.ones(k))
    mu = rng.standard_normal((k, d)) * 3
    cov = np.empty((k, d, d))
    for j in range(k):
        a = rng.standard_normal((d, d))
        cov[j] = a @ a.T / d + 0.3 * np.eye(d)
    return pis, mu, cov


@pytest.mark.parametrize("valu", [False, True])
@pytest.mark.parametrize("d,k", [(16, 8), (2, 3), (7, 33), (12, 20), (16, 40)])
def test_gmm_zstep_vs_oracle(ops, oracle, monkeypatch, d, k, valu):
    """z from the GPU against the oracle's FP64 form with the same uniforms:
    at least 99.9 % equal (the rest within rounding of a bin edge).  A padded
    dimension of 16 (d = 12, 16) runs on the matrix cores unless
    AUGUR_GMM_Z_VALU=1 selects the FP32 vector form; 4000 points leave a
    partial 64-point group."""
    monkeypatch.setenv("AUGUR_GMM_Z_VALU", "1" if valu else "0")
    rng = np.random.default_rng(d * k)
    pis, mu, cov = _gmm_params(rng, k, d)
    n = 4000
    zt = rng.choice(k, n, p=pis)
    y = np.array([rng.multivariate_normal(mu[j], cov[j]) for j in zt]).astype(np.float32)
    z = ops.gmm_zstep(y, pis, mu, cov, 3, 1, 8)
    g = oracle.Gmm(y.astype(np.float64), np.ones(k), np.zeros(d), np.eye(d), d + 2, np.eye(d))
    zr = g.zstep(pis, mu, cov, 3, 1, 8)
    agree = (z == zr).mean()
    assert agree >= 0.999, agree


def test_gmm_conj_vs_oracle(ops, oracle):
    rng = np.random.default_rng(11)
    k, d, n = 5, 4, 3000
    pis, mu, cov = _gmm_params(rng, k, d)
    y = rng.standard_normal((n, d)) * 2 + 1
    z = rng.integers(0, k, n)
    z[z == 4] = 3  # component 4 empty
    cnt = np.bincount(z, minlength=k).astype(np.int32)
    sum_y = np.array([y[z == j].sum(0) for j in range(k)])
    sum_yy = np.array([y[z == j].T @ y[z == j] for j in range(k)])
    alpha, mu0, cov0, df, scale0 = np.ones(k), np.zeros(d), 10 * np.eye(d), d + 2, np.eye(d)
    mu_c = mu.copy()  # current mu: the centre of the statistics
    pis_g, mu_g, cov_g = ops.gmm_conj(alpha, mu0, cov0, df, scale0, cnt, sum_y, sum_yy,
                                      pis, mu, cov, 21, 5)
    pis_r = oracle.dirichlet_conj(21, 0, 5, oracle.TAG_GMM_PI, 0, alpha, cnt)
    np.testing.assert_allclose(pis_g, pis_r, rtol=1e-10)
    for j in range(k):
        mu_r = oracle.mvnormal_conj(21, 0, 5, j * d, mu0, cov0, sum_y[j], cnt[j], cov[j])
        np.testing.assert_allclose(mu_g[j], mu_r, rtol=1e-5, atol=1e-5)
        m = mu_r
        obs = sum_yy[j] - np.outer(m, sum_y[j]) - np.outer(sum_y[j], m) + cnt[j] * np.outer(m, m)
        cov_r = oracle.invwishart(21, 0, 5, j * d * d, df + cnt[j], obs + scale0)
        np.testing.assert_allclose(cov_g[j], cov_r, rtol=2e-4, atol=1e-6)
    del mu_c
