"""The gradient surrogate: its training kernels against the FP64 oracle, the
exactness of surrogate HMC (same posterior as exact HMC), what the trained
network gains, and the fallback when it is poor."""
import numpy as np
import pytest

from augurv2_amd import diagnostics
from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from augurv2_amd import _lib, ops, runtime
    _lib.require_gpu()
    return ops, runtime


def _flat(p):
    return np.concatenate([p["W1"].ravel(), p["b1"], p["W2"].ravel(), p["b2"], p["W3"].ravel(),
                           p["b3"]]).astype(np.float32)


def _unflat(v, d, h):
    v = np.asarray(v, np.float64)
    o = 0
    out = {}
    for k, shape in (("W1", (h, d)), ("b1", (h,)), ("W2", (h, h)), ("b2", (h,)), ("W3", (d, h)),
                     ("b3", (d,))):
        n = int(np.prod(shape))
        out[k] = v[o:o + n].reshape(shape)
        o += n
    return out


@pytest.mark.parametrize("d,h,rows", [(16, 48, 300), (128, 256, 4096)])
def test_mlp_loss_and_gradient_match_oracle(gpu, d, h, rows):
    """Ragged tiles (16, 48, 300) and the benchmark's network on one
    training batch (128, 256, 4096 rows: the weight gradients split k)."""
    ops, _ = gpu
    rng = np.random.default_rng(0)
    p = omlp.init_params(d, h, rng)
    for k in ("b1", "b2", "b3"):
        p[k] = 0.1 * rng.standard_normal(p[k].shape)
    x = rng.standard_normal((rows, d)).astype(np.float32)
    t = rng.standard_normal((rows, d)).astype(np.float32)
    p32 = _unflat(_flat(p), d, h)  # the FP32-rounded parameters the GPU sees
    out, loss, g = ops.mlp_loss_grad(_flat(p), x, t, d, h)
    # the device's products take BF16 operands with FP32 accumulation: the
    # restatement rounds the same operands.  What remains is FP32 rounding
    # (accumulation order, FP32 activations), plus the rare operand whose FP32
    # and FP64 values round to neighbouring BF16 values (one BF16 step, 2^-8,
    # in one term of a sum): hence 2e-3 of each tensor's scale.
    rnd = omlp.bf16_round
    ref_out, _ = omlp.forward(p32, x.astype(np.float64), rnd)
    ref_loss, ref_g = omlp.loss_and_grads(p32, x.astype(np.float64), t.astype(np.float64), rnd)
    assert np.abs(out - ref_out).max() <= 2e-3 * np.abs(ref_out).max()
    assert abs(loss - ref_loss) <= 1e-4 * ref_loss
    gg = _unflat(g, d, h)
    for k in ref_g:
        scale = np.abs(ref_g[k]).max()
        assert np.abs(gg[k] - ref_g[k]).max() <= 2e-3 * scale + 1e-7, k
    # against the exact (unrounded) network the BF16 products are a
    # percent-level approximation, no more
    ex_loss, ex_g = omlp.loss_and_grads(p32, x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - ex_loss) <= 2e-2 * ex_loss
    for k in ex_g:
        assert np.abs(gg[k] - ex_g[k]).max() <= 5e-2 * np.abs(ex_g[k]).max(), k


def test_mlp_adam_step_matches_oracle_and_training_converges(gpu, oracle):
    ops, _ = gpu
    rng = np.random.default_rng(1)
    d, h, n, batch, lr = 8, 32, 2000, 64, 3e-3
    x = rng.standard_normal((n, d)).astype(np.float32)
    w = rng.standard_normal((d, d)) / np.sqrt(d)
    t = np.tanh(x @ w).astype(np.float32)  # a smooth target the network can fit
    p0 = omlp.init_params(d, h, rng)
    p1, hist = ops.mlp_train(_flat(p0), x, t, d, h, steps=1, batch=batch, lr=lr, seed=5)
    # the same minibatch the device drew: row floor(u32 * n / 2^32) of Philox
    # (seed, chain 0) at counter (row, 0, step 0, tag 900)
    idx = [int((int(oracle.draw4(5, 0, r, 0, 0, 900)[0]) * n) >> 32) for r in range(batch)]
    q = _unflat(_flat(p0), d, h)
    loss, g = omlp.loss_and_grads(q, x[idx].astype(np.float64), t[idx].astype(np.float64),
                                  omlp.bf16_round)
    assert abs(hist[0] - loss) <= 1e-4 * loss
    m = {k: np.zeros_like(v) for k, v in q.items()}
    v = {k: np.zeros_like(v) for k, v in q.items()}
    omlp.adam_step(q, g, m, v, 1, lr)
    np.testing.assert_allclose(p1, _flat(q), rtol=0, atol=lr * 1e-2)
    _, hist = ops.mlp_train(_flat(p0), x, t, d, h, steps=600, batch=batch, lr=lr, seed=6)
    assert hist[-20:].mean() < 0.1 * hist[:20].mean()


def _hlr_data(rng, n, d):
    x = rng.standard_normal((n, d)).astype(np.float32)
    beta = rng.standard_normal(d) * 0.4
    y = (rng.random(n) < 1 / (1 + np.exp(-(x @ beta)))).astype(np.int32)
    return x, y


def _run(rt, x, y, surrogate, seed, n_sweeps, train=300, hidden=64, fallback_cfg=None, eps=0.015,
         v_mwg=0.0):
    from augurv2_amd._lib import HMC_BLOCK_THETA, HMC_BLOCK_V, MWG_BLOCK_V
    C, SIM, EPS = 4, 0.15, eps  # 0.15 / eps jittered leapfrog steps
    d = x.shape[1]
    # v by HMC, or by MWG [v] ~ Normal(v, v_mwg) when v_mwg > 0
    vblk = (MWG_BLOCK_V, 0.0, v_mwg, 0) if v_mwg > 0 else (HMC_BLOCK_V, 1.0, 0.05, 0)
    s = rt.HlrSampler(x, y, [vblk, (HMC_BLOCK_THETA, SIM, EPS, 0)],
                      chains=C, seed=seed, ref_compat=False)
    for c in range(C):
        s.set("state", np.concatenate([np.zeros(d), [0.0], [1.0]]), chain=c)
    # exact warmup first: training pairs must come from the typical set, not
    # from the burn-in path, whose gradients are orders of magnitude larger
    s.step(200)
    if surrogate:
        kw = dict(hidden=hidden, train_sweeps=train, max_pairs=200_000, epochs=40, batch=256,
                  lr=2e-3, min_accept=0.6, window=50, seed=3)
        kw.update(fallback_cfg or {})
        s.enable_surrogate(1, **kw)
        s.step(train)  # collection with exact HMC, then training
    s.step(100)
    s.trace("state", n_sweeps)
    s.step(n_sweeps)
    draws = s.read_trace("state")
    info = s.surrogate_info() if surrogate else None
    acc = s.acceptance()
    s.close()
    return draws, info, acc


def _compare(exact, surr, d, n):
    """Posterior means of surrogate HMC against exact HMC: each within 4
    MCSE (family-wise over the d + 1 parameters) and at most 5% (or one)
    beyond 3 MCSE; returns the minimum bulk ESS of each run."""
    keep = list(range(d)) + [d + 1]
    beyond3 = 0
    for j in keep:
        a, b = exact[:, :, j].T, surr[:, :, j].T
        assert diagnostics.rhat(b) < 1.05
        se = np.hypot(diagnostics.mcse_mean(a), diagnostics.mcse_mean(b))
        z = abs(a.mean() - b.mean()) / se
        assert z <= 4.0, (j, a.mean(), b.mean(), se)
        beyond3 += z > 3.0
    assert beyond3 <= max(1, 0.05 * len(keep)), beyond3
    ess_e = min(diagnostics.ess_bulk(exact[:, :, j].T) for j in keep)
    ess_s = min(diagnostics.ess_bulk(surr[:, :, j].T) for j in keep)
    return ess_e, ess_s


@pytest.mark.parametrize("d,hidden,n_rows,eps,v_mwg", [(16, 64, 20000, 0.015, 0.0),
                                                        (64, 128, 20000, 0.015, 0.0),
                                                        (128, 256, 30000, 0.0075, 0.002)])
def test_surrogate_hmc_is_exact_and_pays_off(gpu, d, hidden, n_rows, eps, v_mwg):
    """d = 16: the generic trajectory kernel and FP64 log-density; d = 64,
    hidden = 128: the register-resident BF16-dot trajectory kernel and the
    FP64 matrix-core log-density; d = 128, hidden = 256: the trajectory
    kernel of the N = 1M, D = 128 benchmark (layer 2 in LDS), with half the
    step size (at 0.015 exact HMC itself accepts under half its proposals
    there) and v by MWG (HMC on v accepts ~0.2% there and v does not mix in
    2000 sweeps)."""
    _, rt = gpu
    rng = np.random.default_rng(2)
    x, y = _hlr_data(rng, n_rows, d)
    n = 2000
    exact, _, acc_e = _run(rt, x, y, False, 11, n, eps=eps, v_mwg=v_mwg)
    surr, info, acc_s = _run(rt, x, y, True, 12, n, hidden=hidden, eps=eps, v_mwg=v_mwg)
    assert info["phase"] == 2 and info["fallbacks"] == 0, (info, acc_e, acc_s)
    assert info["window_accept"] >= 0.6, info
    # near the posterior the logistic gradient is close to linear in theta
    assert 0.0 < info["resid_frac"] < 0.2, info
    ess_e, ess_s = _compare(exact, surr, d, n)
    # the gain: effective draws per exact data pass.  Exact HMC spends about
    # 0.15 / eps passes per sweep (the jittered leapfrog count); surrogate HMC
    # one (the log-density of the proposal).
    assert ess_s >= 3.0 * ess_e * eps / 0.15, (ess_s, ess_e, info)


def test_failing_surrogate_trajectories_fall_back_to_exact_hmc(gpu):
    _, rt = gpu
    rng = np.random.default_rng(3)
    x, y = _hlr_data(rng, 5000, 16)
    # fixed surrogate steps of 0.2: eps * sqrt(largest curvature ~ 750) ~ 5.5,
    # far past the leapfrog's stability limit of 2, so proposals are rejected
    draws, info, acc = _run(rt, x, y, True, 13, 300, train=20,
                            fallback_cfg=dict(epochs=1, steps=20, step_size=0.2))
    assert info["phase"] == 3 and info["fallbacks"] == 1, info
    assert info["window_accept"] < 0.6
    assert np.isfinite(draws).all()
    assert acc[1].mean() > 0.3  # exact HMC took over


def test_set_state_after_training_refreshes_the_exact_density(gpu):
    """The surrogate block's Metropolis test uses the FP64 data log-density of
    the current state.  After set_pt moves every chain 3 posterior standard
    deviations off the mean (a state of far lower density), proposals back
    toward the mean must be accepted: judged against the density of the
    state before the move, they would all be rejected."""
    from augurv2_amd import _lib
    _, rt = gpu
    rng = np.random.default_rng(5)
    d = 16
    x, y = _hlr_data(rng, 5000, d)
    from augurv2_amd._lib import HMC_BLOCK_THETA, HMC_BLOCK_V
    C = 4
    s = rt.HlrSampler(x, y, [(HMC_BLOCK_V, 1.0, 0.05, 0), (HMC_BLOCK_THETA, 0.15, 0.015, 0)],
                      chains=C, seed=21, ref_compat=False)
    s.set("state", np.concatenate([np.zeros(d), [0.0], [1.0]]))
    s.step(200)
    s.enable_surrogate(1, hidden=64, train_sweeps=200, max_pairs=100_000, epochs=20, batch=256,
                       lr=2e-3, min_accept=0.0, window=1000, seed=3)
    s.step(200)
    s.trace("state", 300)
    s.step(300)
    draws = s.read_trace("state")[:, :, :d].reshape(-1, d)
    assert s.surrogate_info()["phase"] == 2
    mean, sd = draws.mean(axis=0), draws.std(axis=0)
    for c in range(C):
        st = s.get("state", c)
        st[:d] = mean + 3.0 * sd
        s.set("state", st, chain=c)

    def counts():
        out = np.zeros((2, C, 2))
        _lib.call("augur_stats", s._h, _lib.ptr(out), out.nbytes)
        return out[1]

    before = counts()
    s.step(5)
    acc = counts() - before
    s.close()
    assert acc[:, 1].sum() == 5 * C
    assert acc[:, 0].sum() >= 0.5 * 5 * C, acc


def test_network_gains_over_the_linear_map_on_a_skewed_posterior(gpu):
    """What the trained network adds to the least-squares linear term: on a
    small, strongly informative dataset (N = 300, D = 8, coefficients of
    scale 1.5) the posterior is far from Gaussian and the gradient far from
    linear in theta.  The same surrogate trajectories (0.6 / 0.06 jittered
    leapfrog steps, exact FP64 Metropolis test) then accept far more often
    with the network on the linear map's residual than with the linear map
    alone, and mix faster.  (On the N = 1M benchmark posterior the linear
    term already leaves only ~1e-5 of the variance and the two accept
    alike.)"""
    _, rt = gpu
    from augurv2_amd._lib import HMC_BLOCK_THETA, MWG_BLOCK_V
    rng = np.random.default_rng(1)
    n, d = 300, 8
    x = rng.standard_normal((n, d)).astype(np.float32)
    beta = rng.standard_normal(d) * 1.5
    y = (rng.random(n) < 1 / (1 + np.exp(-(x @ beta)))).astype(np.int32)
    out = {}
    for linear_only in (True, False):
        s = rt.HlrSampler(x, y, [(MWG_BLOCK_V, 0.0, 0.5, 0), (HMC_BLOCK_THETA, 0.6, 0.06, 0)],
                          chains=4, seed=5, ref_compat=False)
        for c in range(4):
            s.set("state", np.concatenate([np.zeros(d), [0.0], [1.0]]), chain=c)
        s.step(300)
        s.enable_surrogate(1, hidden=64, train_sweeps=300, max_pairs=400_000, epochs=60, batch=256,
                           lr=2e-3, min_accept=0.0, window=100_000, seed=3, linear_only=linear_only)
        s.step(300)
        before = np.zeros((2, 4, 2))
        from augurv2_amd import _lib
        _lib.call("augur_stats", s._h, _lib.ptr(before), before.nbytes)
        s.trace("state", 1500)
        s.step(1500)
        after = np.zeros((2, 4, 2))
        _lib.call("augur_stats", s._h, _lib.ptr(after), after.nbytes)
        info = s.surrogate_info()
        draws = s.read_trace("state")
        s.close()
        assert info["phase"] == 2, info
        acc = (after[1, :, 0] - before[1, :, 0]).sum() / (after[1, :, 1] - before[1, :, 1]).sum()
        ess = min(diagnostics.ess_bulk(draws[:, :, j].T) for j in range(d))
        out[linear_only] = (acc, ess, info)
    (acc_l, ess_l, info_l), (acc_n, ess_n, info_n) = out[True], out[False]
    assert info_n["loss"] < 0.5, info_n  # the network explains most of the linear map's residual
    assert acc_n >= 0.85 and acc_n >= acc_l + 0.2, (acc_n, acc_l)
    assert ess_n > ess_l, (ess_n, ess_l)
