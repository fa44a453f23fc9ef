"""Oracle for the batched HLR sampler's linear gradient surrogate (test
infrastructure).

Only tests use this module, as the checker of the device's surrogate sweeps;
it is never the thing measured or shipped.

hmc_step restates one HMC update of a dataset's theta block in which the
leapfrog trajectory takes the data gradient from a given function instead
of the data, while everything else is the exact step of
cbits/augur_mcmclib.h:343-424 as oracle.Hlr.hmc_step (or_hmc_step) follows
it: momentum p0_j ~ N(0, 1) from Philox counter (j, 0, sweep, tag), step
size jittered by U(0.8, 1.2) from counter (0, 0, sweep, tag + 1) when the
step count is given negative, half step / full steps / half step, kinetic
energies 0.5 |p|^2, and the Metropolis test log u < [log p(theta') -
log p(theta)] + K(p0) - K(p) with u from counter (0, 0, sweep, tag + 2),
log p the block's full conditional (data + Normal(0, v) prior) evaluated
exactly.  tag = 16 + 4 * block_id; the batched theta block is block 1.

linear_grad builds the device's surrogate: g(theta) = gbar + W (theta - m)
with W the FP32 values the device holds and theta - m rounded to FP32
before the product (the device forms the product in FP32).

Pinning: with the exact data gradient as grad_fn, hmc_step reproduces
oracle.Hlr.hmc_step (tests/test_oracle_bsurr.py), which fixes the streams,
the leapfrog order and the Metropolis rule; the prior gradient is the
Normal(0, v) dotpt -theta / v of cbits/augur_dist.c:383-394.
"""
import math

import numpy as np

from . import oracle as orc
from .genmodel import jitter, normal_from, u01


def linear_grad(w, mean_theta, mean_grad):
    """The device's linear surrogate of the data gradient."""
    w64 = np.asarray(w, np.float32).astype(np.float64)
    m = np.asarray(mean_theta, np.float64)
    gbar = np.asarray(mean_grad, np.float64)

    def g(theta):
        dv = (np.asarray(theta, np.float64) - m).astype(np.float32).astype(np.float64)
        return gbar + w64 @ dv

    return g


def hmc_step(model, state, grad_fn, eps, fixed_steps, seed, chain, sweep, block_id=1):
    """One surrogate-trajectory HMC update of theta = state[:d] in place
    (state = [theta, b, v] of an oracle.Hlr model without intercept);
    returns whether the proposal was accepted."""
    d = model.d
    v = float(state[d + 1])
    tag = 16 + 4 * block_id
    step, n_steps = jitter(seed, chain, sweep, tag + 1, 0.0, eps, fixed_steps)
    p0 = np.empty(d)
    for j in range(d):
        w = orc.draw4(seed, chain, j, 0, sweep, tag)
        p0[j] = normal_from(int(w[0]), int(w[1]))

    def full(theta):
        return grad_fn(theta) + (-theta / v)

    th = np.array(state[:d], np.float64)
    p = p0 + 0.5 * step * full(th)
    for l in range(n_steps):
        th = th + step * p
        coef = step if l != n_steps - 1 else 0.5 * step
        p = p + coef * full(th)
    k0 = 0.5 * float(np.sum(p0 * p0))
    k1 = 0.5 * float(np.sum(p * p))
    prop = np.array(state, np.float64)
    prop[:d] = th
    lr = (model.block_ll(prop, 2) - model.block_ll(state, 2)) + (k0 - k1)
    u = u01(int(orc.draw4(seed, chain, 0, 0, sweep, tag + 2)[0]))
    acc = math.log(u) < lr
    if acc:
        state[:d] = th
    return acc
