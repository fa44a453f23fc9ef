"""Sampler-level parity through the C-ABI: one sweep of each family on the
GPU against the same sweep of the CPU FP64 restatement, started from the
same state with the same Philox streams."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    from augurv2_amd import _lib, runtime
    _lib.require_gpu()
    return runtime


def _corpus(rng, D, V, mean_len):
    lens = np.maximum(rng.poisson(mean_len, D), 1)
    doc_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    w = rng.integers(0, V, doc_off[-1]).astype(np.int32)
    return doc_off, w


def test_lda_two_sweeps_match_oracle(rt, oracle):
    rng = np.random.default_rng(0)
    K, D, V, C = 6, 40, 120, 2
    doc_off, w = _corpus(rng, D, V, 60)
    alpha, beta = np.full(K, 0.1), np.full(V, 0.01)
    s = rt.LdaSampler(K, doc_off, w, alpha, beta, chains=C, seed=77)
    z0 = [rng.integers(0, K, len(w)).astype(np.int32) for _ in range(C)]
    for c in range(C):
        s.set("z", z0[c], chain=c)
    for sweep in range(2):
        s.step(1)
        for c in range(C):
            theta = np.zeros((D, K))
            phi = np.zeros((K, V))
            z = z0[c].copy()
            oracle.lda_sweep(doc_off, w, alpha, beta, 77, c, sweep, theta, phi, z)
            np.testing.assert_allclose(s.get("theta", c), theta, rtol=1e-9, atol=1e-300)
            np.testing.assert_allclose(s.get("phi", c), phi, rtol=1e-9, atol=1e-300)
            zg = s.get("z", c)
            assert (zg == z).mean() >= 0.999
            z0[c] = zg  # continue both from the GPU's assignment
    ll = s.log_density()
    assert np.isfinite(ll).all()
    s.close()


@pytest.mark.parametrize("K", [12, 20])
def test_lda_ragged_documents_match_oracle(rt, oracle, K):
    """Documents of 0, 1, 63, 64, 65, 128, 129 and 300 tokens (the z step
    takes them in 64-token chunks) and V = 512, two full 256-word blocks of
    each phi row: three sweeps of two chains against the oracle, each sweep
    continuing from the GPU's assignment, so the count tables the chunks add
    up are checked through the next sweep's theta and phi.  K = 12 takes the
    phi normalisation inside the z step, K = 20 the Dirichlet kernel's
    last-block normalisation."""
    rng = np.random.default_rng(21)
    lens = np.array([0, 1, 63, 64, 65, 128, 129, 300, 0, 5, 64, 0])
    D, V, C = len(lens), 512, 2
    doc_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    w = rng.integers(0, V, int(doc_off[-1])).astype(np.int32)
    alpha, beta = np.full(K, 0.3), np.full(V, 0.02)
    s = rt.LdaSampler(K, doc_off, w, alpha, beta, chains=C, seed=5)
    z0 = [rng.integers(0, K, len(w)).astype(np.int32) for _ in range(C)]
    for c in range(C):
        s.set("z", z0[c], chain=c)
    for sweep in range(3):
        s.step(1)
        for c in range(C):
            theta, phi, z = np.zeros((D, K)), np.zeros((K, V)), z0[c].copy()
            oracle.lda_sweep(doc_off, w, alpha, beta, 5, c, sweep, theta, phi, z)
            np.testing.assert_allclose(s.get("theta", c), theta, rtol=1e-9, atol=1e-300)
            np.testing.assert_allclose(s.get("phi", c), phi, rtol=1e-9, atol=1e-300)
            zg = s.get("z", c)
            assert (zg == z).mean() >= 0.995
            z0[c] = zg
    s.close()


def test_lda_traced_phi_is_normalised_every_sweep(rt, oracle):
    """With K <= 16 the normalised phi is written after the sweeps of a step,
    and after every sweep while a trace records.  The first traced phi equals
    the oracle's sweep from the same z; every traced phi is a set of
    distributions and differs from the sweep before (no stale copy); the
    last equals what get returns."""
    rng = np.random.default_rng(8)
    K, D, V = 9, 30, 200
    doc_off, w = _corpus(rng, D, V, 70)
    alpha, beta = np.full(K, 0.2), np.full(V, 0.03)
    s = rt.LdaSampler(K, doc_off, w, alpha, beta, chains=1, seed=4)
    z = rng.integers(0, K, len(w)).astype(np.int32)
    s.set("z", z)
    s.trace("phi", 3)
    s.step(3)
    tr = s.read_trace("phi").reshape(3, K, V)
    theta, phi = np.zeros((D, K)), np.zeros((K, V))
    oracle.lda_sweep(doc_off, w, alpha, beta, 4, 0, 0, theta, phi, z.copy())
    np.testing.assert_allclose(tr[0], phi, rtol=1e-9, atol=1e-300)
    np.testing.assert_allclose(tr.sum(axis=2), 1.0, rtol=1e-12)
    assert not np.array_equal(tr[0], tr[1]) and not np.array_equal(tr[1], tr[2])
    np.testing.assert_array_equal(s.get("phi").reshape(K, V), tr[2])
    s.close()


def test_lda_log_density_is_the_model_density(rt, oracle):
    """augur_get_curr_ll for lda.rv is the reference's ll_mod: Dirichlet
    terms of theta and phi without normaliser (the oracle's dirichlet_ll,
    pinned against scipy) plus log theta[d, z] + log phi[z, w] per token."""
    rng = np.random.default_rng(6)
    K, D, V, C = 7, 25, 90, 2
    doc_off, w = _corpus(rng, D, V, 50)
    alpha, beta = rng.uniform(0.05, 0.5, K), rng.uniform(0.01, 0.2, V)
    s = rt.LdaSampler(K, doc_off, w, alpha, beta, chains=C, seed=5)
    for c in range(C):
        s.set("z", rng.integers(0, K, len(w)).astype(np.int32), chain=c)
    s.step(2)
    ll = s.log_density()
    doc = np.repeat(np.arange(D), np.diff(doc_off))
    for c in range(C):
        theta, phi, z = s.get("theta", c), s.get("phi", c), s.get("z", c)
        ref = sum(oracle.dirichlet_ll(theta[d], alpha) for d in range(D))
        ref += sum(oracle.dirichlet_ll(phi[k], beta) for k in range(K))
        ref += np.sum(np.log(theta[doc, z])) + np.sum(np.log(phi[z, w]))
        assert abs(ll[c] - ref) <= 1e-10 * abs(ref), (c, ll[c], ref)
    s.close()


def test_lda_counts_exact_after_set(rt, oracle):
    """Counts are recomputed from a set z; the Dirichlet draws then consume
    them, so equal theta/phi to the oracle imply equal count tables."""
    rng = np.random.default_rng(3)
    K, D, V = 12, 30, 64
    doc_off, w = _corpus(rng, D, V, 40)
    alpha, beta = np.full(K, 0.2), np.full(V, 0.05)
    s = rt.LdaSampler(K, doc_off, w, alpha, beta, seed=1)
    z = rng.integers(0, K, len(w)).astype(np.int32)
    s.set("z", z)
    s.step(1)
    theta = np.zeros((D, K))
    phi = np.zeros((K, V))
    zz = z.copy()
    oracle.lda_sweep(doc_off, w, alpha, beta, 1, 0, 0, theta, phi, zz)
    np.testing.assert_allclose(s.get("theta"), theta, rtol=1e-9)
    np.testing.assert_allclose(s.get("phi"), phi, rtol=1e-9)
    with pytest.raises(Exception):
        s.set("z", np.full(len(w), K, np.int32))  # out of range topic


def _hlr_data(rng, n, d):
    x = rng.standard_normal((n, d)).astype(np.float32)
    beta = rng.standard_normal(d) * 0.5
    y = (rng.random(n) < 1 / (1 + np.exp(-(x @ beta)))).astype(np.int32)
    return x, y


@pytest.mark.parametrize("fixed", [0, 12])
@pytest.mark.parametrize("intercept", [False, True])
def test_hlr_iterations_match_oracle(rt, oracle, fixed, intercept):
    from augurv2_amd._lib import HMC_BLOCK_V, HMC_BLOCK_B, HMC_BLOCK_THETA
    rng = np.random.default_rng(10 + fixed)
    n, d, C = 3000, 8, 4
    x, y = _hlr_data(rng, n, d)
    blocks = [(HMC_BLOCK_V, 1.0, 0.05, fixed)]
    if intercept:
        blocks.append((HMC_BLOCK_B, 0.1, 0.01, fixed))
    blocks.append((HMC_BLOCK_THETA, 0.1, 0.01, fixed))
    s = rt.HlrSampler(x, y, blocks, chains=C, seed=123, has_intercept=intercept)
    m = oracle.Hlr(x.astype(np.float64), y, has_intercept=intercept)
    states = []
    for c in range(C):
        st = np.concatenate([rng.standard_normal(d) * 0.3, [0.1 * intercept], [0.8 + 0.1 * c]])
        s.set("state", st, chain=c)
        states.append(st)
    kinds = {HMC_BLOCK_V: 0, HMC_BLOCK_B: 1, HMC_BLOCK_THETA: 2}
    for sweep in range(3):
        s.step(1)
        for c in range(C):
            for bi, (kind, sl, eps, fx) in enumerate(blocks):
                m.hmc_step(states[c], kinds[kind], bi, sl, eps, fx, 123, c, sweep)
            np.testing.assert_allclose(s.get("state", c), states[c], rtol=2e-5, atol=1e-7)
    acc = s.acceptance()
    assert acc.shape == (len(blocks), C)
    s.close()


@pytest.mark.parametrize("d,C,intercept", [(64, 4, False), (64, 4, True), (128, 8, False), (32, 2, True)])
def test_hlr_gradient_only_passes_give_the_same_trajectories(rt, monkeypatch, d, C, intercept):
    """The register-form passes before the shortest chain's last leapfrog
    step skip the log-density (the intercept's gradient sum is kept).  The
    adjoints are the same values, so every state must equal, bit for bit,
    a run whose passes all compute the log-density (AUGUR_HLR_GRAD_PASS=0);
    the step jitter gives the chains different step counts."""
    from augurv2_amd._lib import HMC_BLOCK_V, HMC_BLOCK_B, HMC_BLOCK_THETA
    rng = np.random.default_rng(40 + d)
    x, y = _hlr_data(rng, 3000, d)
    blocks = [(HMC_BLOCK_V, 1.0, 0.05, 0)]
    if intercept:
        blocks.append((HMC_BLOCK_B, 0.1, 0.01, 0))
    blocks.append((HMC_BLOCK_THETA, 0.1, 0.005, 0))
    starts = [np.concatenate([rng.standard_normal(d) * 0.1, [0.1 * intercept], [0.8 + 0.1 * c]]) for c in range(C)]
    runs = []
    for grad_pass in ("1", "0"):
        monkeypatch.setenv("AUGUR_HLR_GRAD_PASS", grad_pass)
        s = rt.HlrSampler(x, y, blocks, chains=C, seed=77, has_intercept=intercept)
        for c in range(C):
            s.set("state", starts[c], chain=c)
        s.step(5)
        runs.append([s.get("state", c).copy() for c in range(C)])
        acc = s.acceptance()
        s.close()
    assert acc[-1].mean() > 0.3
    for a, b in zip(*runs):
        np.testing.assert_array_equal(a, b)


def test_hlr_v_block_last_matches_oracle(rt, oracle):
    """HMC [theta] (*) HMC [v]: a v block with no data block after it runs as
    its own launch instead of sharing the next block's first launch."""
    from augurv2_amd._lib import HMC_BLOCK_V, HMC_BLOCK_THETA
    rng = np.random.default_rng(31)
    n, d, C = 2500, 8, 4
    x, y = _hlr_data(rng, n, d)
    blocks = [(HMC_BLOCK_THETA, 0.1, 0.01, 0), (HMC_BLOCK_V, 1.0, 0.05, 0)]
    s = rt.HlrSampler(x, y, blocks, chains=C, seed=77)
    m = oracle.Hlr(x.astype(np.float64), y)
    states = []
    for c in range(C):
        st = np.concatenate([rng.standard_normal(d) * 0.3, [0.0], [0.7 + 0.1 * c]])
        s.set("state", st, chain=c)
        states.append(st)
    for sweep in range(3):
        s.step(1)
        for c in range(C):
            m.hmc_step(states[c], 2, 0, 0.1, 0.01, 0, 77, c, sweep)
            m.hmc_step(states[c], 0, 1, 1.0, 0.05, 0, 77, c, sweep)
            np.testing.assert_allclose(s.get("state", c), states[c], rtol=2e-5, atol=1e-7)
    s.close()


@pytest.mark.parametrize("ref_compat", [False, True])
def test_hlr_mwg_v_matches_oracle(rt, oracle, ref_compat):
    """MWG [v] ~ Normal(v, 0.3) (*) HMC [theta]: every sweep's state against
    the oracle, starting from v values where the proposal often goes below 0."""
    from augurv2_amd._lib import HMC_BLOCK_THETA, MWG_BLOCK_V
    rng = np.random.default_rng(21)
    n, d, C = 2000, 8, 4
    x, y = _hlr_data(rng, n, d)
    blocks = [(MWG_BLOCK_V, 0.0, 0.3, 0), (HMC_BLOCK_THETA, 0.1, 0.01, 0)]
    s = rt.HlrSampler(x, y, blocks, chains=C, seed=99, ref_compat=ref_compat)
    m = oracle.Hlr(x.astype(np.float64), y, ref_compat=ref_compat)
    states = []
    for c in range(C):
        st = np.concatenate([rng.standard_normal(d) * 0.3, [0.0], [0.2 + 0.3 * c]])
        s.set("state", st, chain=c)
        states.append(st)
    moved = 0
    for sweep in range(6):
        s.step(1)
        for c in range(C):
            v_before = states[c][d + 1]
            m.mwg_v_step(states[c], 0, 0.3, 99, c, sweep)
            moved += states[c][d + 1] != v_before
            m.hmc_step(states[c], 2, 1, 0.1, 0.01, 0, 99, c, sweep)
            np.testing.assert_allclose(s.get("state", c), states[c], rtol=2e-5, atol=1e-7)
    assert moved > 0
    s.close()


@pytest.mark.parametrize("ref_compat", [False, True])
def test_hlr_mwg_b_matches_oracle(rt, oracle, ref_compat):
    """compiler/augur/test/hlr2.infer: MWG [v] ~ Normal(v, 1.0) (*) MWG [b] ~
    Normal(b, s) (*) HMC [theta] (with a proposal variance the data accept
    sometimes), every sweep's state against the oracle."""
    from augurv2_amd._lib import HMC_BLOCK_THETA, MWG_BLOCK_B, MWG_BLOCK_V
    rng = np.random.default_rng(22)
    n, d, C = 2000, 8, 4
    x, y = _hlr_data(rng, n, d)
    y = ((x @ (rng.standard_normal(d) * 0.3) + 0.7 + rng.logistic(size=n)) > 0).astype(np.int32)
    blocks = [(MWG_BLOCK_V, 0.0, 1.0, 0), (MWG_BLOCK_B, 0.0, 0.01, 0), (HMC_BLOCK_THETA, 0.1, 0.01, 0)]
    s = rt.HlrSampler(x, y, blocks, chains=C, seed=98, has_intercept=True, ref_compat=ref_compat)
    m = oracle.Hlr(x.astype(np.float64), y, has_intercept=True, ref_compat=ref_compat)
    states = []
    for c in range(C):
        st = np.concatenate([rng.standard_normal(d) * 0.3, [0.2 * c], [1.0 + 0.3 * c]])
        s.set("state", st, chain=c)
        states.append(st)
    moved = 0
    for sweep in range(8):
        s.step(1)
        for c in range(C):
            m.mwg_v_step(states[c], 0, 1.0, 98, c, sweep)
            b_before = states[c][d]
            m.mwg_b_step(states[c], 1, 0.01, 98, c, sweep)
            moved += states[c][d] != b_before
            m.hmc_step(states[c], 2, 2, 0.1, 0.01, 0, 98, c, sweep)
            np.testing.assert_allclose(s.get("state", c), states[c], rtol=2e-5, atol=1e-7)
    assert moved > 0
    acc = s.acceptance()
    assert np.all(acc[1] > 0.0)
    # the deterministic sigmoid(x . theta + b) of dethlr.rv, read per chain
    for c in range(C):
        st = s.get("state", c)
        ref = 1.0 / (1.0 + np.exp(-(x.astype(np.float64) @ st[:d] + st[d])))
        np.testing.assert_allclose(s.get("bias", c), ref, rtol=1e-12)
    s.close()


def test_hlr_curr_ll(rt, oracle):
    from augurv2_amd._lib import HMC_BLOCK_V, HMC_BLOCK_THETA
    rng = np.random.default_rng(4)
    x, y = _hlr_data(rng, 2000, 6)
    s = rt.HlrSampler(x, y, [(HMC_BLOCK_V, 1.0, 0.05, 0), (HMC_BLOCK_THETA, 0.1, 0.01, 0)],
                      chains=2, seed=3)
    m = oracle.Hlr(x.astype(np.float64), y)
    st = np.concatenate([rng.standard_normal(6) * 0.2, [0.0], [1.3]])
    s.set("state", st)
    ll = s.log_density()
    ref = m.block_ll(st, 0) + m.data_ll(st)  # v prior + theta priors + data
    np.testing.assert_allclose(ll, ref, rtol=1e-6)


def _gmm_setup(rng, n, d, k):
    pis = rng.dirichlet(np.ones(k))
    mu = rng.standard_normal((k, d)) * 3
    zt = rng.choice(k, n, p=pis)
    y = (mu[zt] + rng.standard_normal((n, d))).astype(np.float32)
    return y, zt


@pytest.mark.parametrize("sample_pi,sample_cov", [(True, True), (False, False)])
def test_gmm_log_density_matches_oracle(rt, oracle, sample_pi, sample_cov):
    """augur_get_curr_ll is the model's log density (the reference's ll_mod):
    priors of pis / mu / cov and the categorical and data terms, at a set
    state and after sweeps.  Data terms in FP32 per point (2e-6 relative)."""
    from scipy.stats import invwishart
    rng = np.random.default_rng(21)
    n, d, k, C = 5000, 5, 7, 2
    y, _ = _gmm_setup(rng, n, d, k)
    alpha, mu0, cov0, df, scale0 = rng.uniform(0.5, 2, k), rng.standard_normal(d), 6 * np.eye(d), d + 3, np.eye(d)
    s = rt.GmmSampler(y, alpha, mu0, cov0, df, scale0, chains=C, seed=4, sample_pi=sample_pi,
                      sample_cov=sample_cov)
    g = oracle.Gmm(y.astype(np.float64), alpha, mu0, cov0, df, scale0, sample_pi, sample_cov)
    for c in range(C):
        s.set("pis", rng.dirichlet(np.ones(k)), chain=c)
        s.set("mu", rng.standard_normal((k, d)) * 2, chain=c)
        s.set("cov", np.stack([invwishart.rvs(df, scale0, random_state=rng) for _ in range(k)]), chain=c)
        s.set("z", rng.integers(0, k, n).astype(np.int32), chain=c)
    for step in (0, 2):
        s.step(step)
        ll = s.log_density()
        for c in range(C):
            ref = g.ll_mod(s.get("pis", c), s.get("mu", c), s.get("cov", c), s.get("z", c))
            assert abs(ll[c] - ref) <= 2e-6 * abs(ref), (step, c, ll[c], ref)
    s.close()


def test_gmm_draws_do_not_depend_on_the_z_grid(rt, monkeypatch):
    """The matrix-core z step (D = 16) run with 1, 7 and the default number of
    blocks per chain: every draw comes from its own counter, so four sweeps
    leave bit-identical states."""
    rng = np.random.default_rng(12)
    n, d, k, C = 20000, 16, 8, 2
    y, _ = _gmm_setup(rng, n, d, k)
    init = [(rng.standard_normal((k, d)) * 2, rng.integers(0, k, n).astype(np.int32)) for _ in range(C)]
    out = []
    for blocks in ("1", "7", None):
        if blocks is None:
            monkeypatch.delenv("AUGUR_GMM_ZBLOCKS", raising=False)
        else:
            monkeypatch.setenv("AUGUR_GMM_ZBLOCKS", blocks)
        s = rt.GmmSampler(y, np.ones(k), np.zeros(d), 10 * np.eye(d), d + 2, np.eye(d), chains=C,
                          seed=6)
        for c, (mu, z) in enumerate(init):
            s.set("mu", mu, chain=c)
            s.set("z", z, chain=c)
        s.step(4)
        out.append([[s.get(name, c) for name in ("pis", "mu", "cov", "z")] for c in range(C)])
        s.close()
    for other in out[1:]:
        for c in range(C):
            for a, b in zip(out[0][c], other[c]):
                np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("d,k,n,sample_pi,sample_cov", [
    (4, 3, 3000, True, True), (4, 3, 3000, False, False),
    # the benchmark's shape (padded dimension 16, 32 components: the sorted
    # matrix-core statistics kernel, the 16 x 16 FP64 conjugate draws and
    # the matrix-core z step)
    (16, 32, 6000, True, True),
    (16, 20, 4000, True, True),   # 20 of 32 padded components live
    (12, 10, 3000, True, True),   # 12 of 16 padded dimensions: the sorted kernel's zero columns
    (16, 40, 4000, True, True),   # 64 padded components: the one-hot statistics kernel
    (8, 40, 4000, True, True),    # padded dimension 8, 64 padded components
])
def test_gmm_sweeps_match_oracle(rt, oracle, d, k, n, sample_pi, sample_cov):
    """Whole sweeps (statistics, conjugate draws of pis / mu / cov, z step)
    against the oracle's sweep (cbits/augur_dist.c:526-538, 633-669, 740-757;
    Core/CgConj.hs:171-246) on the same Philox draws: pis to 1e-9 relative
    pins the counts, mu and cov the sums and scatter matrices."""
    rng = np.random.default_rng(8 + d + k)
    C = 2
    y, zt = _gmm_setup(rng, n, d, k)
    alpha, mu0, cov0, df, scale0 = np.ones(k), np.zeros(d), 10 * np.eye(d), d + 2, np.eye(d)
    s = rt.GmmSampler(y, alpha, mu0, cov0, df, scale0, chains=C, seed=9, sample_pi=sample_pi,
                      sample_cov=sample_cov)
    g = oracle.Gmm(y.astype(np.float64), alpha, mu0, cov0, df, scale0, sample_pi, sample_cov)
    st = []
    for c in range(C):
        pis = np.full(k, 1.0 / k)
        mu = rng.standard_normal((k, d))
        cov = np.stack([np.eye(d) * (1 + 0.2 * j) for j in range(k)])
        z = rng.integers(0, k, n).astype(np.int32)
        for name, val in (("pis", pis), ("mu", mu), ("cov", cov), ("z", z)):
            s.set(name, val, chain=c)
        st.append([pis, mu, cov, z])
    for sweep in range(2):
        s.step(1)
        for c in range(C):
            pis, mu, cov, z = st[c]
            g.sweep(pis, mu, cov, z, 9, c, sweep)
            np.testing.assert_allclose(s.get("pis", c), pis, rtol=1e-9)
            np.testing.assert_allclose(s.get("mu", c), mu, rtol=1e-5, atol=1e-5)
            np.testing.assert_allclose(s.get("cov", c), cov, rtol=1e-4, atol=1e-6)
            zg = s.get("z", c)
            assert (zg == z).mean() >= 0.999
            st[c][3] = zg.copy()
    s.close()
