"""The experiments harness's measurements over the device sampler: the
predictive log-likelihood of held-out data (experiments/lda/lda.py:149-173,
experiments/hmvgmm/hmvgmm.py:143-178), the compile / sample timing of an
inference run (experiments/util/infersys.py:57-80) and the stats.yml record
(experiments/util/modelenv.py:73-86).

The reference computes the predictive log-likelihood in Python loops with
scipy; here each term is one device reduction (augur_op_lda_predictive_ll,
augur_op_gmm_predictive_ll, FP64) and the running means follow the
reference's bookkeeping.
"""
import ctypes as C
import os
import time

import numpy as np

from . import _lib as L


def _csr(ragged):
    lens = np.array([len(r) for r in ragged], np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    flat = (np.concatenate([np.asarray(r, np.int32) for r in ragged]) if lens.sum()
            else np.zeros(0, np.int32))
    return off, np.ascontiguousarray(flat, np.int32)


def lda_predictive_ll_term(theta, phi, w_test_ho):
    """sum_d sum_j log sum_k theta[d][k] phi[k][w_test_ho[d][j]] (one sample)."""
    th = L.f64(np.stack([np.asarray(t, np.float64) for t in theta]))
    ph = L.f64(np.stack([np.asarray(p, np.float64) for p in phi]))
    off, w = _csr(w_test_ho)
    out = C.c_double(0.0)
    L.call("augur_op_lda_predictive_ll", th.shape[0], th.shape[1], ph.shape[1], L.ptr(th), L.ptr(ph),
           L.ptr(off), L.ptr(w), C.byref(out))
    return out.value


def lda_predictive_ll(hypers, phi, thetas, w_test_ho):
    """experiments/lda/lda.py:149-173: the running mean of the held-out
    log-likelihood over the last 10 samples of theta (phi fixed); returns
    its final value."""
    S = len(thetas)
    prev_ll = 0.0
    pred_ll = np.zeros(S)
    for s in range(max(0, S - 10), S):
        prev_ll += lda_predictive_ll_term(thetas[s], phi, w_test_ho)
        pred_ll[s] = prev_ll / (s - (S - 10) + 1.0)
    return pred_ll[S - 1]


def gmm_predictive_ll_term(pis, mu, cov, test_y):
    """sum_n log sum_k pis[k] N(test_y[n] | mu[k], cov[k]) (one sample)."""
    p = L.f64(pis)
    m = L.f64(np.stack([np.asarray(r, np.float64) for r in mu]))
    cv = L.f64(np.stack([np.asarray(r, np.float64) for r in cov]))
    y = L.f64(test_y)
    out = C.c_double(0.0)
    L.call("augur_op_gmm_predictive_ll", len(p), m.shape[1], L.ptr(p), L.ptr(m), L.ptr(cv), y.shape[0],
           L.ptr(y), C.byref(out))
    return out.value


def gmm_predictive_ll(pis, mus, covs, test_y, init_pt, S):
    """experiments/hmvgmm/hmvgmm.py:143-178: the held-out log-likelihood at
    the initial point, then its running mean over the S samples."""
    init_ll = gmm_predictive_ll_term(init_pt["pis"], init_pt["mu"], init_pt["cov"], test_y)
    prev_ll = 0.0
    pred_ll = np.zeros(S)
    for s in range(S):
        prev_ll += gmm_predictive_ll_term(pis[s], mus[s], covs[s], test_y)
        pred_ll[s] = prev_ll / (s + 1.0)
    return np.append([init_ll], pred_ll)


class Timed:
    """compile and sample times of an AugurInfer run, measured as
    experiments/util/infersys.py does around compile(...)(...) and samplen
    (wall clock here: the sampler runs on the device)."""

    def __init__(self):
        self.compile_time = 0.0
        self.sample_time = 0.0

    def compile(self, infer, hypers, data):
        t0 = time.perf_counter()
        infer.compile(*hypers)(*data)
        self.compile_time = time.perf_counter() - t0

    def samplen(self, infer, **kw):
        t0 = time.perf_counter()
        out = infer.samplen(**kw)
        self.sample_time = time.perf_counter() - t0
        return out


def dump_pred_ll_stats(path_to_stat, compile_time, sample_time, pred_ll):
    """experiments/util/modelenv.py:73-86: <path_to_stat>/stats.yml with the
    compile and sample times and the predictive log-likelihood list."""
    os.makedirs(path_to_stat, exist_ok=True)
    path = os.path.join(path_to_stat, "stats.yml")
    with open(path, "w") as f:
        f.write("compile: " + str(compile_time) + "\n")
        f.write("sample: " + str(sample_time) + "\n")
        f.write("predll: " + "[" + ", ".join(map(str, pred_ll)) + "]" + "\n")
    return path
