"""The batched surrogate's oracle step (oracle/bsurr.py) against the exact
HMC restatement: given the exact data gradient it must be the exact step,
which pins its Philox streams, leapfrog order and Metropolis rule."""
import numpy as np
import pytest


@pytest.mark.parametrize("fixed_steps", [8, -8])
def test_surrogate_step_with_the_exact_gradient_is_the_exact_step(oracle, fixed_steps):
    from oracle import bsurr
    rng = np.random.default_rng(3)
    n, d = 400, 16
    x = rng.standard_normal((n, d))
    beta = rng.standard_normal(d) * 0.3
    y = (rng.random(n) < 1 / (1 + np.exp(-(x @ beta)))).astype(np.int32)
    m = oracle.Hlr(x, y, prior_kind=0, prior_a=1.0, prior_b=1.0, ref_compat=False)

    def exact(theta):
        return m.data_grad(np.concatenate([theta, [0.0, 1.0]]))[0]

    outcomes = set()
    for sweep in range(12):
        st0 = np.concatenate([beta + rng.standard_normal(d) * 0.05, [0.0], [0.5 + 0.1 * sweep]])
        a, b = st0.copy(), st0.copy()
        m.hmc_step(a, 2, 1, 0.0, 0.06, fixed_steps, 11, 3, sweep)
        acc = bsurr.hmc_step(m, b, exact, 0.06, fixed_steps, 11, 3, sweep)
        np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-14, err_msg=f"sweep {sweep}")
        outcomes.add(bool(acc))
    assert outcomes == {True, False}  # both branches of the Metropolis test were exercised


def test_linear_surrogate_rounds_like_the_device(oracle):
    """theta - m enters the product rounded to FP32 and W as FP32 values; with
    an exactly representable case the map is the affine function itself."""
    from oracle import bsurr
    w = np.array([[2.0, -1.0], [0.5, 4.0]])
    g = bsurr.linear_grad(w, [1.0, -2.0], [3.0, 5.0])
    np.testing.assert_array_equal(g(np.array([2.0, 0.0])), [3.0 + 2.0 - 2.0, 5.0 + 0.5 + 8.0])
    third = float(np.float32(1.0 / 3.0))  # 0.3333333432674408: theta - m after FP32 rounding
    assert third != 1.0 / 3.0
    np.testing.assert_array_equal(g(np.array([1.0 + 1.0 / 3.0, -2.0])), [3.0 + 2.0 * third, 5.0 + 0.5 * third])
