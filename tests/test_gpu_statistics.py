"""Posterior parity: the GPU sampler and the CPU FP64 restatement of the
reference sampler, run as independent chains on the same small problems,
agree on posterior means and variances within 3 Monte Carlo standard
errors (combined).  Seeds are fixed, so each test is deterministic."""
import numpy as np
import pytest

from augurv2_amd import diagnostics

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    from augurv2_amd import _lib, runtime
    _lib.require_gpu()
    return runtime


def _agree(a, b, k=3.0):
    """a, b: (chains, draws) of one scalar; means and variances within k
    combined MCSE."""
    ma, mb = a.mean(), b.mean()
    se = np.hypot(diagnostics.mcse_mean(a), diagnostics.mcse_mean(b))
    assert abs(ma - mb) <= k * se, (ma, mb, se)
    va, vb = (a - ma) ** 2, (b - mb) ** 2
    sev = np.hypot(diagnostics.mcse_mean(va), diagnostics.mcse_mean(vb))
    assert abs(va.mean() - vb.mean()) <= k * sev, (va.mean(), vb.mean(), sev)


def test_hlr_posterior_matches_oracle(rt, oracle):
    from augurv2_amd._lib import HMC_BLOCK_B, HMC_BLOCK_THETA, HMC_BLOCK_V
    rng = np.random.default_rng(21)
    n, d, C, iters, burn = 400, 3, 4, 3000, 300
    x = rng.standard_normal((n, d)).astype(np.float32)
    y = (rng.random(n) < 1 / (1 + np.exp(-(x @ np.array([1.0, -0.5, 0.25]) + 0.3)))).astype(np.int32)
    blocks = [(HMC_BLOCK_V, 1.0, 0.05, 0), (HMC_BLOCK_B, 0.3, 0.05, 0),
              (HMC_BLOCK_THETA, 0.3, 0.05, 0)]
    # textbook Normal density: the reference's normal_dotarg2 makes HMC on v
    # accept rarely, and its slow mixing leaves too few effective draws
    s = rt.HlrSampler(x, y, blocks, chains=C, seed=5, prior_kind=1, prior_a=1.0,
                      has_intercept=True, ref_compat=False)
    init = [np.array([0.1 * c, -0.1, 0.0, 0.0, 1.0]) for c in range(C)]
    for c in range(C):
        s.set("state", init[c], chain=c)
    s.step(burn)
    s.trace("state", iters)
    s.step(iters)
    g = s.read_trace("state")  # (iters, C, d + 2)
    s.close()
    m = oracle.Hlr(x.astype(np.float64), y, prior_kind=1, prior_a=1.0, has_intercept=True,
                   ref_compat=False)
    kinds = {HMC_BLOCK_V: 0, HMC_BLOCK_B: 1, HMC_BLOCK_THETA: 2}
    o = np.zeros_like(g)
    for c in range(C):
        st = init[c].copy()
        for it in range(burn + iters):
            for bi, (k, sl, eps, fx) in enumerate(blocks):
                m.hmc_step(st, kinds[k], bi, sl, eps, fx, 99, c, it)  # independent streams
            if it >= burn:
                o[it - burn, c] = st
    for j in range(d + 2):
        assert diagnostics.rhat(g[:, :, j].T) < 1.05
        _agree(g[:, :, j].T, o[:, :, j].T)


def test_hlr_reference_target_posterior_matches_oracle(rt, oracle):
    """The reference's own target (ref_compat: normal_ll = -log var - t^2 /
    (2 var), its normal_dotarg2 and MWG's extra Hastings term, cbits/augur_dist.c
    :372-394, Core/CgUserProp.hs:130-141) under the schedule of
    examples/hlr.py on hlr.rv: MWG [v] ~ Normal(v, 1.0) (*) HMC [b] (*) HMC
    [theta], v ~ Exponential(1).  GPU chains and independent oracle chains
    agree on posterior means and variances within 3 MCSE."""
    from augurv2_amd._lib import HMC_BLOCK_B, HMC_BLOCK_THETA, MWG_BLOCK_V
    rng = np.random.default_rng(23)
    n, d, C, iters, burn = 400, 3, 4, 3000, 300
    x = rng.standard_normal((n, d)).astype(np.float32)
    y = (rng.random(n) < 1 / (1 + np.exp(-(x @ np.array([1.0, -0.5, 0.25]) + 0.3)))).astype(np.int32)
    blocks = [(MWG_BLOCK_V, 0.0, 1.0, 0), (HMC_BLOCK_B, 0.3, 0.05, 0), (HMC_BLOCK_THETA, 0.3, 0.05, 0)]
    s = rt.HlrSampler(x, y, blocks, chains=C, seed=6, prior_kind=1, prior_a=1.0,
                      has_intercept=True, ref_compat=True)
    init = [np.array([0.1 * c, -0.1, 0.0, 0.0, 1.0]) for c in range(C)]
    for c in range(C):
        s.set("state", init[c], chain=c)
    s.step(burn)
    s.trace("state", iters)
    s.step(iters)
    g = s.read_trace("state")
    acc_v = s.acceptance()[0]
    s.close()
    assert 0.05 < np.nanmean(acc_v) < 0.95, acc_v  # v moves
    m = oracle.Hlr(x.astype(np.float64), y, prior_kind=1, prior_a=1.0, has_intercept=True,
                   ref_compat=True)
    o = np.zeros_like(g)
    for c in range(C):
        st = init[c].copy()
        for it in range(burn + iters):
            m.mwg_v_step(st, 0, 1.0, 98, c, it)  # independent streams
            m.hmc_step(st, 1, 1, 0.3, 0.05, 0, 98, c, it)
            m.hmc_step(st, 2, 2, 0.3, 0.05, 0, 98, c, it)
            if it >= burn:
                o[it - burn, c] = st
    for j in range(d + 2):
        assert diagnostics.rhat(g[:, :, j].T) < 1.05
        assert diagnostics.rhat(o[:, :, j].T) < 1.05
        _agree(g[:, :, j].T, o[:, :, j].T)


def test_lda_posterior_matches_oracle(rt, oracle):
    """Label-invariant summaries: per-document topic concentration
    sum_k theta_dk^2 and the probability that two fixed tokens share a topic."""
    rng = np.random.default_rng(4)
    K, D, V = 3, 6, 12
    lens = rng.integers(8, 15, D)
    doc_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    w = rng.integers(0, V, doc_off[-1]).astype(np.int32)
    alpha, beta = np.full(K, 0.5), np.full(V, 0.2)
    C, iters, burn = 4, 3000, 200
    s = rt.LdaSampler(K, doc_off, w, alpha, beta, chains=C, seed=8)
    for c in range(C):
        s.set("z", rng.integers(0, K, len(w)).astype(np.int32), chain=c)
    s.step(burn)
    s.trace("theta", iters)
    s.trace("z", iters)
    s.step(iters)
    th = s.read_trace("theta")  # (iters, C, D, K)
    zz = s.read_trace("z")      # (iters, C, T)
    s.close()
    t1, t2 = 0, int(doc_off[1])  # first token of documents 0 and 1
    g_conc = (th**2).sum(-1)     # (iters, C, D)
    g_same = (zz[:, :, t1] == zz[:, :, t2]).astype(float)
    o_conc = np.zeros_like(g_conc)
    o_same = np.zeros_like(g_same)
    for c in range(C):
        theta, phi = np.zeros((D, K)), np.zeros((K, V))
        z = rng.integers(0, K, len(w)).astype(np.int32)
        for it in range(burn + iters):
            oracle.lda_sweep(doc_off, w, alpha, beta, 31, c, it, theta, phi, z)
            if it >= burn:
                o_conc[it - burn, c] = (theta**2).sum(-1)
                o_same[it - burn, c] = z[t1] == z[t2]
    for dd in range(D):
        _agree(g_conc[:, :, dd].T, o_conc[:, :, dd].T)
    _agree(g_same.T, o_same.T)


def test_gmm_posterior_matches_oracle(rt, oracle):
    """Two well-separated clusters in 2-D (no label switching within the
    run): posterior of the sorted component means and of pis."""
    rng = np.random.default_rng(6)
    n, d, k = 300, 2, 2
    zt = rng.random(n) < 0.35
    y = np.where(zt[:, None], np.array([3.0, 0.0]), np.array([-3.0, 1.0])) + \
        rng.standard_normal((n, d)) * 0.8
    alpha, mu0, cov0, df, scale0 = np.ones(k), np.zeros(d), 10 * np.eye(d), d + 2, np.eye(d)
    C, iters, burn = 4, 2000, 100
    s = rt.GmmSampler(y.astype(np.float32), alpha, mu0, cov0, df, scale0, chains=C, seed=12)
    starts = []
    for c in range(C):
        st = dict(pis=np.full(k, 0.5), mu=np.array([[2.0, 0.0], [-2.0, 0.0]]),
                  cov=np.stack([np.eye(d)] * k), z=(y[:, 0] < 0).astype(np.int32))
        for name, val in st.items():
            s.set(name, val, chain=c)
        starts.append(st)
    s.step(burn)
    for name in ("mu", "pis", "cov"):
        s.trace(name, iters)
    s.step(iters)
    gm, gp, gc = s.read_trace("mu"), s.read_trace("pis"), s.read_trace("cov")
    s.close()
    g = oracle.Gmm(y.astype(np.float32).astype(np.float64), alpha, mu0, cov0, df, scale0)
    om, op, oc = np.zeros_like(gm), np.zeros_like(gp), np.zeros_like(gc)
    for c in range(C):
        st = {kk: np.array(v, copy=True) for kk, v in starts[c].items()}
        for it in range(burn + iters):
            g.sweep(st["pis"], st["mu"], st["cov"], st["z"], 77, c, it)
            if it >= burn:
                om[it - burn, c], op[it - burn, c], oc[it - burn, c] = st["mu"], st["pis"], st["cov"]
    for j in range(k):
        _agree(gp[:, :, j].T, op[:, :, j].T)
        for a in range(d):
            _agree(gm[:, :, j, a].T, om[:, :, j, a].T)
            _agree(gc[:, :, j, a, a].T, oc[:, :, j, a, a].T)
