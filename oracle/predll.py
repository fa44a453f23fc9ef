"""The experiments harness's predictive log-likelihood terms in numpy/scipy.

TEST INFRASTRUCTURE ONLY (the checker of augur_op_*_predictive_ll): the
reference's loops, experiments/lda/lda.py:149-173 (inner sum over held-out
tokens) and experiments/hmvgmm/hmvgmm.py:143-178 (inner_ll), with
scipy.special.logsumexp in place of the deprecated scipy.misc one.
"""
import numpy as np
from scipy.special import logsumexp
from scipy.stats import multivariate_normal


def lda_term(theta, phi, w_test_ho):
    K = phi.shape[0]
    acc = 0.0
    tmp = np.zeros(K)
    for d in range(len(w_test_ho)):
        for j in range(len(w_test_ho[d])):
            for k in range(K):
                tmp[k] = np.log(theta[d][k]) + np.log(phi[k][w_test_ho[d][j]])
            acc += logsumexp(tmp)
    return acc


def gmm_term(pis, mu, cov, test_y):
    K = len(pis)
    acc = 0.0
    tmp = np.zeros(K)
    for n in range(len(test_y)):
        for k in range(K):
            tmp[k] = multivariate_normal.logpdf(test_y[n], mu[k], cov[k]) + np.log(pis[k])
        acc += logsumexp(tmp)
    return acc
