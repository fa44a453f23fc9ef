"""Batched HLR (many datasets of uneven size and width in one launch):
parity of every dataset's chains with the CPU restatement run on the same
BF16-rounded X, graph replay equal to direct launches, and the autotuner."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    from augurv2_amd import _lib, runtime
    _lib.require_gpu()
    return runtime


def bf16_round(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _datasets(rng, shapes):
    xs, ys = [], []
    for n, d in shapes:
        x = rng.standard_normal((n, d)).astype(np.float32)
        beta = rng.standard_normal(d) * 0.3
        y = (rng.random(n) < 1 / (1 + np.exp(-(x @ beta)))).astype(np.int32)
        xs.append(x)
        ys.append(y)
    return xs, ys


SHAPES = [(300, 16), (1000, 32), (2500, 64), (777, 16), (130, 128), (64, 32)]


def _init(s, rng, C):
    starts = {}
    for i, (n, d) in enumerate(SHAPES):
        for c in range(C):
            st = np.concatenate([rng.standard_normal(d) * 0.1, [0.0], [1.0 + 0.1 * i + 0.05 * c]])
            s.set(f"state:{i}", st, chain=c)
            starts[i, c] = st
    return starts


@pytest.mark.parametrize("C,form", [(2, "register"), (2, "lds"), (4, "register_mfma"),
                                    (4, "register"), (3, "register_mfma"), (1, "register")])
@pytest.mark.parametrize("L", [8, -8])
def test_batched_matches_oracle(rt, oracle, monkeypatch, L, C, form):
    """L > 0: fixed step and count; L < 0: the reference's step jitter with
    |L| steps (every chain runs the same number of passes).  The register
    form of the pass (at 3-4 chains with its a x phase on the matrix cores,
    or all on the vector units when forced), and the LDS-staged form
    forced.  HMC on v with steps of 0.01: at 0.05 its leapfrog is unstable
    on dataset 3 and the oracle alone turns a 1e-7 nudge of theta into 28
    times this test's tolerance within three sweeps; at 0.01 the same nudge
    stays below a quarter of it on every dataset."""
    monkeypatch.setenv("AUGUR_BHLR_PASS_LDS", "1" if form == "lds" else "0")
    monkeypatch.setenv("AUGUR_BHLR_R_VALU", "1" if form == "register" else "0")
    rng = np.random.default_rng(0)
    xs, ys = _datasets(rng, SHAPES)
    EPS = 0.01
    s = rt.BhlrSampler(xs, ys, chains=C, seed=7, v_block=(0.2, 0.01, 0), theta_block=(0.1, EPS, L),
                       rows_per_block=512, use_graph=True)
    starts = _init(s, rng, C)
    models = [oracle.Hlr(bf16_round(x).astype(np.float64), y, prior_kind=0, prior_a=1.0,
                         prior_b=1.0, ref_compat=False) for x, y in zip(xs, ys)]
    for sweep in range(3):
        s.step(1)
        for i in range(len(SHAPES)):
            for c in range(C):
                st = starts[i, c]
                gc = i * C + c
                models[i].hmc_step(st, 0, 0, 0.2, 0.01, 0, 7, gc, sweep)
                models[i].hmc_step(st, 2, 1, 0.1, EPS, L, 7, gc, sweep)
                np.testing.assert_allclose(s.get(f"state:{i}", c), st, rtol=2e-5, atol=1e-7)
    acc = s.acceptance()
    assert acc.shape == (2, C, len(SHAPES)) and acc[1].mean() > 0.3
    s.close()


@pytest.mark.parametrize("vector", [False, True])
def test_batched_matrix_core_pass_matches_oracle(rt, oracle, monkeypatch, vector):
    """12 chains per dataset (padded to 16 matrix-core columns): BF16 GEMM
    form of the data pass for D = 16, 32, 64, 128, against the oracle; the
    vector form on the same chains for comparison.

    v moves by MWG here.  With HMC on v (20 jittered steps of 0.05) the
    oracle alone turns a 1e-7 relative nudge of the start into a different
    accept decision on dataset 4 (N = 130, D = 128) within two sweeps: that
    leapfrog is unstable at v's curvature, so such a case compares rounding
    luck, not kernels.  With MWG the same nudge stays below 2e-5."""
    monkeypatch.setenv("AUGUR_BHLR_VECTOR", "1" if vector else "0")
    rng = np.random.default_rng(8)
    xs, ys = _datasets(rng, SHAPES)
    C, L, EPS = 12, -6, 0.01
    s = rt.BhlrSampler(xs, ys, chains=C, seed=23, v_block=("mwg", 0.5), theta_block=(0.1, EPS, L),
                       rows_per_block=512, use_graph=False)
    starts = _init(s, rng, C)
    models = [oracle.Hlr(bf16_round(x).astype(np.float64), y, prior_kind=0, prior_a=1.0,
                         prior_b=1.0, ref_compat=False) for x, y in zip(xs, ys)]
    for sweep in range(3):
        s.step(1)
        for i in range(len(SHAPES)):
            for c in range(C):
                st = starts[i, c]
                gc = i * C + c
                models[i].mwg_v_step(st, 0, 0.5, 23, gc, sweep)
                models[i].hmc_step(st, 2, 1, 0.1, EPS, L, 23, gc, sweep)
                np.testing.assert_allclose(s.get(f"state:{i}", c), st, rtol=2e-5, atol=1e-7,
                                           err_msg=f"dataset {i} chain {c} sweep {sweep}")
    assert s.acceptance()[1].mean() > 0.3
    s.close()


def test_batched_mwg_v_matches_oracle(rt, oracle):
    """v block by Metropolis-within-Gibbs, proposal Normal(v, 0.5), with the
    step-jittered theta block, graph replay on."""
    rng = np.random.default_rng(5)
    xs, ys = _datasets(rng, SHAPES)
    C, L, EPS = 2, -6, 0.01
    s = rt.BhlrSampler(xs, ys, chains=C, seed=17, v_block=("mwg", 0.5), theta_block=(0.1, EPS, L),
                       rows_per_block=512, use_graph=True)
    starts = _init(s, rng, C)
    models = [oracle.Hlr(bf16_round(x).astype(np.float64), y, prior_kind=0, prior_a=1.0,
                         prior_b=1.0, ref_compat=False) for x, y in zip(xs, ys)]
    for sweep in range(3):
        s.step(1)
        for i in range(len(SHAPES)):
            for c in range(C):
                st = starts[i, c]
                gc = i * C + c
                models[i].mwg_v_step(st, 0, 0.5, 17, gc, sweep)
                models[i].hmc_step(st, 2, 1, 0.1, EPS, L, 17, gc, sweep)
                np.testing.assert_allclose(s.get(f"state:{i}", c), st, rtol=2e-5, atol=1e-7)
    assert s.acceptance()[0].mean() > 0.05
    s.close()


def test_batched_warmup_adaptation(rt, oracle):
    """Per-dataset step sizes (dual averaging) and MWG proposal variances
    (Robbins-Monro) adapt to each dataset, acceptance lands near the target,
    and the frozen sampler is plain MCMC: a later sweep equals the oracle's
    with the adapted per-dataset settings."""
    rng = np.random.default_rng(11)
    xs, ys = _datasets(rng, SHAPES)
    C, L = 4, -10
    s = rt.BhlrSampler(xs, ys, chains=C, seed=41, v_block=("mwg", 1.0), theta_block=(0.1, 0.01, L),
                       rows_per_block=512, use_graph=True)
    _init(s, rng, C)
    s.adapt(300, 0.8)
    eps, vv = s.step_sizes()
    assert np.all(np.isfinite(eps)) and np.all(eps > 0) and eps.max() / eps.min() > 2.0
    assert np.all(np.isfinite(vv)) and np.all(vv > 0)
    s.step(200)
    acc = s.acceptance()[1]  # includes the warmup's sweeps
    assert np.all((acc.mean(axis=0) > 0.5) & (acc.mean(axis=0) < 0.98)), acc.mean(axis=0)
    models = [oracle.Hlr(bf16_round(x).astype(np.float64), y, prior_kind=0, prior_a=1.0,
                         prior_b=1.0, ref_compat=False) for x, y in zip(xs, ys)]
    sweep = s.sweeps()
    states = {(i, c): s.get(f"state:{i}", c).copy() for i in range(len(SHAPES)) for c in range(C)}
    s.step(1)
    for i in range(len(SHAPES)):
        for c in range(C):
            st = states[i, c]
            gc = i * C + c
            models[i].mwg_v_step(st, 0, vv[i], 41, gc, sweep)
            models[i].hmc_step(st, 2, 1, 0.1, eps[i], L, 41, gc, sweep)
            # absolute 1e-6: the FP32 gradient over N rows is exact to about
            # 1e-7 sum|a x| (~1e-4 at N = 2500), and ten adapted steps move
            # theta by eps^2 L times that, ~3e-7 at eps = 0.02
            np.testing.assert_allclose(s.get(f"state:{i}", c), st, rtol=2e-5, atol=1e-6,
                                       err_msg=f"dataset {i} chain {c}")
    s.close()


def test_graph_replay_equals_direct_launches(rt):
    rng = np.random.default_rng(1)
    xs, ys = _datasets(rng, SHAPES)
    out = []
    for graph in (False, True):
        s = rt.BhlrSampler(xs, ys, chains=4, seed=3, theta_block=(0.1, 0.01, 6), rows_per_block=256,
                           use_graph=graph)
        _init(s, np.random.default_rng(9), 4)
        s.step(2)
        s.trace("state", 3)
        s.step(3)
        out.append(s.read_trace("state"))
        s.close()
    assert np.array_equal(out[0], out[1])


def test_autotuner_picks_a_timed_candidate(rt):
    rng = np.random.default_rng(2)
    xs, ys = _datasets(rng, [(20000, 64), (5000, 16), (60000, 32)])
    s = rt.BhlrSampler(xs, ys, chains=4, seed=1, rows_per_block=0)
    rows, table = s.tuning()
    assert rows in table and table[rows] == min(table.values())
    assert all(v > 0 for v in table.values())
    s.step(2)
    assert np.isfinite(s.log_density()).all()
    s.close()


def _surrogate_sampler(rt, rng, C, L, seed, train, **kw):
    xs, ys = _datasets(rng, SHAPES)
    s = rt.BhlrSampler(xs, ys, chains=C, seed=seed, v_block=("mwg", 1.0), theta_block=(0.1, 0.01, L),
                       rows_per_block=512, use_graph=True)
    _init(s, rng, C)
    s.adapt(100, 0.8)
    s.surrogate(train_sweeps=train, **kw)
    return s, xs, ys


def test_batched_surrogate_fit_is_least_squares_on_the_pairs(rt):
    """Each dataset's map is the least-squares fit of its captured (theta,
    data gradient) pairs: means of theta and of the gradient, then the
    regression on standardised theta with 1e-6 added to the correlation
    matrix's diagonal, as numpy solves it in FP64 (W kept in FP32)."""
    rng = np.random.default_rng(5)
    C, L, train = 2, -8, 30
    s, _, _ = _surrogate_sampler(rt, rng, C, L, 3, train, ridge=1e-6, min_accept=0.0, window=0)
    assert s.surrogate_info()["phase"] == 1
    s.step(train)
    info = s.surrogate_info()
    assert info["phase"] == 2 and info["datasets_on_surrogate"] == len(SHAPES)
    assert info["pairs_per_dataset"] == train * 8 * C
    for i, (n, d) in enumerate(SHAPES):
        t, g = s.surrogate_pairs(i)
        assert t.shape == (train * 8 * C, d) and np.isfinite(t).all() and np.isfinite(g).all()
        w, m, gbar = s.surrogate_map(i)
        t64, g64 = t.astype(np.float64), g.astype(np.float64)
        np.testing.assert_allclose(m, t64.mean(0), rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(gbar, g64.mean(0), rtol=1e-12, atol=1e-12 * np.abs(g64).max())
        dt, dg = t64 - m, g64 - gbar
        sd = np.sqrt((dt ** 2).mean(0))
        sd[sd == 0] = 1.0
        z = dt / sd
        r = z.T @ z / len(z) + 1e-6 * np.eye(d)
        b = np.linalg.solve(r, z.T @ dg / len(z))
        w_ref = (b / sd[:, None]).T
        np.testing.assert_allclose(w, w_ref, rtol=1e-4, atol=1e-5 * np.abs(w_ref).max(),
                                   err_msg=f"dataset {i}")
        # the map explains the pairs' gradients (the posterior is near Gaussian)
        resid = dg - dt @ w.T
        assert (resid ** 2).sum() < 0.1 * (dg ** 2).sum(), i
    s.close()


def test_batched_surrogate_sweeps_match_oracle(rt, oracle):
    """Surrogate sweeps (one trajectory kernel, one exact pass at the
    proposal, the FP64 Metropolis test) equal the oracle's step with the
    device's fitted map, dataset by dataset and chain by chain."""
    from oracle import bsurr
    rng = np.random.default_rng(6)
    C, L, seed, train = 2, -8, 9, 20
    s, xs, ys = _surrogate_sampler(rt, rng, C, L, seed, train, ridge=1e-6, min_accept=0.0, window=0)
    s.step(train + 3)
    assert s.surrogate_info()["phase"] == 2
    eps, vv = s.step_sizes()
    maps = [bsurr.linear_grad(*s.surrogate_map(i)) for i in range(len(SHAPES))]
    models = [oracle.Hlr(bf16_round(x).astype(np.float64), y, prior_kind=0, prior_a=1.0,
                         prior_b=1.0, ref_compat=False) for x, y in zip(xs, ys)]
    states = {(i, c): s.get(f"state:{i}", c).copy() for i in range(len(SHAPES)) for c in range(C)}
    accepted = 0
    for k in range(3):
        sweep = s.sweeps()
        s.step(1)
        for i in range(len(SHAPES)):
            for c in range(C):
                st = states[i, c]
                gc = i * C + c
                models[i].mwg_v_step(st, 0, vv[i], seed, gc, sweep)
                accepted += bsurr.hmc_step(models[i], st, maps[i], eps[i], L, seed, gc, sweep)
                np.testing.assert_allclose(s.get(f"state:{i}", c), st, rtol=2e-5, atol=1e-6,
                                           err_msg=f"sweep {k} dataset {i} chain {c}")
    assert accepted >= len(SHAPES) * C  # the surrogate proposals are mostly accepted
    s.close()


@pytest.mark.parametrize("graph", [True, False])
def test_batched_surrogate_mixed_modes_match_oracle(rt, oracle, graph):
    """Some datasets on exact HMC, the rest on the surrogate: the exact
    datasets' leapfrog passes run over their row blocks only, the last pass
    over every block.  Each dataset's sweep equals the oracle's step of its
    own mode (exact HMC or the surrogate trajectory), chain by chain; the
    mode can be switched back."""
    from oracle import bsurr
    rng = np.random.default_rng(16)
    C, L, seed, train = 2, -8, 19, 20
    xs, ys = _datasets(rng, SHAPES)
    s = rt.BhlrSampler(xs, ys, chains=C, seed=seed, v_block=("mwg", 1.0), theta_block=(0.1, 0.01, L),
                       rows_per_block=512, use_graph=graph)
    _init(s, rng, C)
    s.adapt(100, 0.8)
    s.surrogate(train_sweeps=train, ridge=1e-6, min_accept=0.0, window=0)
    s.step(train + 1)
    exact = {1, 2, 5}  # 1000 x 32 (two row blocks), 2500 x 64 (five), 64 x 32 (one)
    for i in exact:
        s.surrogate_set_mode(i, 0)
    modes = s.surrogate_modes()
    assert [int(m) for m in modes] == [0 if i in exact else 1 for i in range(len(SHAPES))]
    assert s.surrogate_info()["datasets_on_surrogate"] == len(SHAPES) - len(exact)
    eps, vv = s.step_sizes()
    maps = [bsurr.linear_grad(*s.surrogate_map(i)) for i in range(len(SHAPES))]
    models = [oracle.Hlr(bf16_round(x).astype(np.float64), y, prior_kind=0, prior_a=1.0,
                         prior_b=1.0, ref_compat=False) for x, y in zip(xs, ys)]
    states = {(i, c): s.get(f"state:{i}", c).copy() for i in range(len(SHAPES)) for c in range(C)}
    for k in range(4):
        if k == 2:  # dataset 2 back on its surrogate
            s.surrogate_set_mode(2, 1)
            exact.discard(2)
        sweep = s.sweeps()
        s.step(1)
        for i in range(len(SHAPES)):
            for c in range(C):
                st = states[i, c]
                gc = i * C + c
                models[i].mwg_v_step(st, 0, vv[i], seed, gc, sweep)
                if i in exact:
                    models[i].hmc_step(st, 2, 1, 0.1, eps[i], L, seed, gc, sweep)
                else:
                    bsurr.hmc_step(models[i], st, maps[i], eps[i], L, seed, gc, sweep)
                np.testing.assert_allclose(s.get(f"state:{i}", c), st, rtol=2e-5, atol=1e-6,
                                           err_msg=f"sweep {k} dataset {i} chain {c} exact {i in exact}")
    with pytest.raises(RuntimeError):
        s.surrogate_set_mode(len(SHAPES), 0)
    s.close()


def test_batched_surrogate_falls_back_to_exact_hmc(rt, oracle):
    """With an acceptance floor no window can meet, every dataset returns to
    exact HMC after one window, and its later sweeps are the exact step."""
    rng = np.random.default_rng(8)
    C, L, seed, train = 2, -8, 13, 10
    s, xs, ys = _surrogate_sampler(rt, rng, C, L, seed, train, ridge=1e-6, min_accept=1.01, window=2)
    s.step(train + 2)
    info = s.surrogate_info()
    assert info["datasets_on_surrogate"] == 0 and info["fallbacks"] == len(SHAPES)
    assert not s.surrogate_modes().any()
    eps, vv = s.step_sizes()
    models = [oracle.Hlr(bf16_round(x).astype(np.float64), y, prior_kind=0, prior_a=1.0,
                         prior_b=1.0, ref_compat=False) for x, y in zip(xs, ys)]
    sweep = s.sweeps()
    states = {(i, c): s.get(f"state:{i}", c).copy() for i in range(len(SHAPES)) for c in range(C)}
    s.step(1)
    for i in range(len(SHAPES)):
        for c in range(C):
            st = states[i, c]
            gc = i * C + c
            models[i].mwg_v_step(st, 0, vv[i], seed, gc, sweep)
            models[i].hmc_step(st, 2, 1, 0.1, eps[i], L, seed, gc, sweep)
            np.testing.assert_allclose(s.get(f"state:{i}", c), st, rtol=2e-5, atol=1e-6,
                                       err_msg=f"dataset {i} chain {c}")
    s.close()


def test_batched_surrogate_posterior_matches_exact_sampler(rt):
    """Posterior means of theta from surrogate sweeps agree with exact
    batched HMC on the same datasets within 4.5 combined Monte Carlo
    standard errors (every coordinate of three datasets), at acceptance
    above 0.6."""
    from augurv2_amd import diagnostics as dg
    rng = np.random.default_rng(12)
    shapes = [(4000, 16), (12000, 32), (8000, 64)]
    xs, ys = _datasets(rng, shapes)
    C, L, n = 4, -10, 1500
    draws = []
    for surrogate in (False, True):
        s = rt.BhlrSampler(xs, ys, chains=C, seed=21 + surrogate, v_block=("mwg", 1.0),
                           theta_block=(0.1, 0.01, L), rows_per_block=512, use_graph=True)
        for i, (nn, d) in enumerate(shapes):
            for c in range(C):
                s.set(f"state:{i}", np.concatenate([rng.standard_normal(d) * 0.05, [0.0, 1.0]]), chain=c)
        s.adapt(400, 0.8)
        if surrogate:
            s.surrogate(train_sweeps=200, ridge=1e-6, min_accept=0.3, window=100)
            s.step(200)
        s.trace("state", n)
        s.step(n)
        tr = s.read_trace("state")  # (n, C, SW)
        if surrogate:
            info = s.surrogate_info()
            assert info["datasets_on_surrogate"] == len(shapes), info
            acc = s.acceptance()[1]
            assert acc.mean() > 0.6, acc
        draws.append(tr)
        s.close()
    off = 0
    worst = 0.0
    for nn, d in shapes:
        for j in range(d):
            xa = draws[0][:, :, off + j].T
            xb = draws[1][:, :, off + j].T
            assert dg.rhat(xa) < 1.05 and dg.rhat(xb) < 1.05
            se = np.hypot(dg.mcse_mean(xa), dg.mcse_mean(xb))
            z = abs(xa.mean() - xb.mean()) / se
            worst = max(worst, z)
        off += d + 2
    assert worst < 4.5, worst


@pytest.mark.parametrize("C", [4, 3, 16])
def test_gradient_only_passes_give_the_same_trajectories(rt, monkeypatch, C):
    """The leapfrog passes before a trajectory's last run kernels that skip
    the log-density (register form with matrix-core a x at 3-4 chains,
    BF16 matrix cores at 16).  Their adjoints are the same values, so every
    state must equal, bit for bit, a run whose passes all compute the
    log-density (AUGUR_BHLR_GRAD_PASS=0)."""
    rng = np.random.default_rng(5)
    xs, ys = _datasets(rng, SHAPES)
    runs = []
    for grad_pass in ("1", "0"):
        monkeypatch.setenv("AUGUR_BHLR_GRAD_PASS", grad_pass)
        s = rt.BhlrSampler(xs, ys, chains=C, seed=29, v_block=("mwg", 0.5), theta_block=(0.1, 0.01, -8),
                           rows_per_block=512, use_graph=True)
        _init(s, np.random.default_rng(6), C)
        s.step(4)
        runs.append([s.get(f"state:{i}", c).copy() for i in range(len(SHAPES)) for c in range(C)])
        acc = s.acceptance()[1]
        s.close()
    assert acc.mean() > 0.3
    for a, b in zip(*runs):
        np.testing.assert_array_equal(a, b)
