"""Writes tests/golden/golden.npz: fixtures that pin the CPU oracle from
sources independent of it.

- Philox4x32-10 known-answer vectors published with Random123
  (Salmon et al., SC'11; kat_vectors file of the Random123 distribution).
- Cholesky factors, inverses and MvNormal log-densities of fixed SPD
  matrices from numpy.linalg / scipy.stats.
- HLR data log-density and gradient at fixed points, in float64 closed
  form (log sigmoid via logaddexp, gradient X^T (y - sigmoid(X beta))).
- The tiny LDA corpus of the reference's examples/lda.py (D = 4, K = 4,
  V = 8, designed topics) as data.

Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.stats as sps

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    g = {}
    g["philox_ctr"] = np.array([[0, 0, 0, 0], [0xffffffff] * 4,
                                [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344]], np.uint32)
    g["philox_key"] = np.array([[0, 0], [0xffffffff] * 2, [0xa4093822, 0x299f31d0]], np.uint32)
    g["philox_out"] = np.array([[0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8],
                                [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd],
                                [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]], np.uint32)

    rng = np.random.default_rng(20170618)
    for n in (2, 5, 16):
        a = rng.standard_normal((n, n))
        spd = a @ a.T + n * np.eye(n)
        g[f"spd{n}"] = spd
        g[f"chol{n}"] = np.linalg.cholesky(spd)
        g[f"inv{n}"] = np.linalg.inv(spd)
        g[f"linv{n}"] = np.linalg.inv(np.linalg.cholesky(spd))
        pt = rng.standard_normal(n)
        mean = rng.standard_normal(n)
        g[f"mvn_pt{n}"] = pt
        g[f"mvn_mean{n}"] = mean
        g[f"mvn_ll{n}"] = np.array(sps.multivariate_normal(mean, spd).logpdf(pt))

    n, d = 200, 7
    x = rng.standard_normal((n, d))
    beta = rng.standard_normal(d) * 2.0
    b = 0.4
    y = (rng.random(n) < 1 / (1 + np.exp(-(x @ beta + b)))).astype(np.int32)
    t = x @ beta + b
    ll = np.sum(np.where(y == 1, -np.logaddexp(0, -t), -np.logaddexp(0, t)))
    s = np.exp(-np.logaddexp(0, -t))
    g.update(hlr_x=x, hlr_y=y, hlr_beta=beta, hlr_b=np.array(b), hlr_ll=np.array(ll),
             hlr_grad=x.T @ (y - s), hlr_grad_b=np.array(np.sum(y - s)))

    # examples/lda.py:41-70 (mk_synthetic_dataset_d4_k4_v8)
    g["lda_w"] = np.array([[0, 2, 1, 3, 0, 2, 1, 3, 0, 2], [4, 6, 5, 7, 4, 6, 5, 7, 5, 7],
                           [2, 4, 3, 5, 2, 4, 3, 5, 2, 4], [0, 6, 1, 7, 0, 6, 1, 7, 1, 7]],
                          np.int32)
    np.savez(os.path.join(HERE, "golden.npz"), **g)


if __name__ == "__main__":
    main()
