"""ctypes wrapper of the CPU FP64 oracle (oracle/augur_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` as the checker.  The product never
imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")

TAG_LDA_THETA, TAG_LDA_PHI, TAG_LDA_Z = 1, 2, 3
TAG_GMM_PI, TAG_GMM_MU, TAG_GMM_COV, TAG_GMM_Z = 4, 5, 6, 7
TAG_HMC_BASE = 16

_lib = None

P_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
P_F = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
P_I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
P_U = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
U32 = C.c_uint32


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE], stdout=subprocess.DEVNULL)


class HlrModel(C.Structure):
    _fields_ = [("n", C.c_int), ("d", C.c_int), ("x", C.c_void_p), ("y", C.c_void_p),
                ("prior_kind", C.c_int), ("prior_a", C.c_double), ("prior_b", C.c_double),
                ("has_intercept", C.c_int), ("ref_compat", C.c_int)]


class GmmModel(C.Structure):
    _fields_ = [("n", C.c_int), ("d", C.c_int), ("k", C.c_int), ("y", C.c_void_p),
                ("alpha", C.c_void_p), ("mu0", C.c_void_p), ("cov0", C.c_void_p),
                ("df", C.c_int), ("scale0", C.c_void_p), ("sample_pi", C.c_int),
                ("sample_cov", C.c_int), ("mu_kernel", C.c_int), ("mu_sim_len", C.c_double),
                ("mu_step_size", C.c_double), ("mu_fixed_steps", C.c_int),
                ("mu_prop_cov", C.c_void_p), ("ref_compat", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        sig = {
            "or_philox4x32_10": (None, [P_U, P_U, P_U]),
            "or_draw4": (None, [U32] * 6 + [P_U]),
            "or_u01": (C.c_double, [U32]),
            "or_u01f": (C.c_float, [U32]),
            "or_normal": (C.c_double, [U32] * 6),
            "or_gamma": (C.c_double, [U32] * 5 + [C.c_double]),
            "or_expit": (C.c_double, [C.c_double]),
            "or_bernoulli_ll": (C.c_double, [C.c_int, C.c_double]),
            "or_bernoulli_dotarg1": (C.c_double, [C.c_int, C.c_double]),
            "or_normal_ll": (C.c_double, [C.c_double] * 3 + [C.c_int]),
            "or_normal_dotpt": (C.c_double, [C.c_double] * 3),
            "or_normal_dotarg2": (C.c_double, [C.c_double] * 3 + [C.c_int]),
            "or_invgamma_ll": (C.c_double, [C.c_double] * 3),
            "or_invgamma_dotpt": (C.c_double, [C.c_double] * 3),
            "or_exponential_ll": (C.c_double, [C.c_double] * 2),
            "or_exponential_dotpt": (C.c_double, [C.c_double] * 2),
            "or_categorical_ref": (C.c_int, [P_D, C.c_int, C.c_double]),
            "or_categorical_f32": (C.c_int, [P_F, P_F, C.c_int, C.c_float]),
            "or_categorical_logits": (C.c_int, [P_D, C.c_int, C.c_double]),
            "or_dirichlet_conj": (None, [U32] * 5 + [P_D, C.c_void_p, C.c_int, P_D]),
            "or_mat_cholesky": (None, [P_D, C.c_int]),
            "or_mat_lower_trmi": (None, [P_D, C.c_int]),
            "or_mat_cholesky_invert": (None, [P_D, C.c_int]),
            "or_mat_trtrmm": (None, [P_D, P_D, C.c_int]),
            "or_mat_trvm": (None, [P_D, P_D, C.c_int]),
            "or_mvnormal_ll": (C.c_double, [P_D, P_D, P_D, C.c_int]),
            "or_mvnormal_mvnormal_conj": (None, [U32] * 4 + [P_D, P_D, P_D, C.c_int, P_D,
                                                             C.c_int, P_D]),
            "or_invwishart_sample": (None, [U32] * 4 + [C.c_int, P_D, C.c_int, P_D]),
            "or_lda_counts": (None, [C.c_int] * 3 + [P_I, P_I, P_I, P_I, P_I]),
            "or_lda_zstep_f32": (None, [C.c_int, C.c_int, P_I, P_I, P_F, P_F, U32, U32, U32,
                                        P_I]),
            "or_lda_sweep": (None, [C.c_int] * 3 + [P_I, P_I, P_D, P_D, U32, U32, U32, P_D,
                                                    P_D, P_I]),
            "or_hlr_data_ll": (C.c_double, [C.POINTER(HlrModel), P_D, C.c_int]),
            "or_hlr_data_grad": (None, [C.POINTER(HlrModel), P_D, P_D, C.POINTER(C.c_double),
                                        C.c_int]),
            "or_hlr_block_ll": (C.c_double, [C.POINTER(HlrModel), P_D, C.c_int]),
            "or_hlr_block_grad": (None, [C.POINTER(HlrModel), P_D, C.c_int, P_D]),
            "or_hmc_step": (C.c_int, [C.POINTER(HlrModel), P_D, C.c_int, C.c_int, C.c_double,
                                      C.c_double, C.c_int, U32, U32, U32, P_D]),
            "or_mwg_v_step": (C.c_int, [C.POINTER(HlrModel), P_D, C.c_int, C.c_double, U32, U32,
                                        U32, P_D]),
            "or_mwg_b_step": (C.c_int, [C.POINTER(HlrModel), P_D, C.c_int, C.c_double, U32, U32,
                                        U32, P_D]),
            "or_gmm_sweep": (None, [C.POINTER(GmmModel), P_D, P_D, P_D, P_I, U32, U32, U32]),
            "or_gmm_zstep": (None, [C.POINTER(GmmModel), P_D, P_D, P_D, P_I, U32, U32, U32]),
            "or_multi_lgamma": (C.c_double, [C.c_int, C.c_double]),
            "or_dirichlet_ll": (C.c_double, [P_D, P_D, C.c_int]),
            "or_invwishart_ll": (C.c_double, [P_D, C.c_int, P_D, C.c_int]),
            "or_gmm_ll_mod": (C.c_double, [C.POINTER(GmmModel), P_D, P_D, P_D, P_I]),
            "or_gmm_comp_like": (C.c_double, [C.POINTER(GmmModel), P_I, C.c_int, P_D, P_D]),
            "or_gmm_mu_fc": (C.c_double, [C.POINTER(GmmModel), P_I, P_D, P_D]),
            "or_gmm_mu_grad": (None, [C.POINTER(GmmModel), P_I, P_D, P_D, P_D]),
            "or_gmm_mu_move": (C.c_int, [C.POINTER(GmmModel), P_D, P_D, P_I, U32, U32, U32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


# ---- thin numpy helpers -------------------------------------------------

def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def draw4(seed, chain, site, sub, sweep, tag):
    out = np.zeros(4, np.uint32)
    lib().or_draw4(seed, chain, site, sub, sweep, tag, out)
    return out


def gamma(seed, chain, site, sweep, tag, shape):
    return lib().or_gamma(seed, chain, site, sweep, tag, float(shape))


def dirichlet_conj(seed, chain, sweep, tag, site0, alpha, cnts):
    alpha = np.ascontiguousarray(alpha, np.float64)
    out = np.zeros(len(alpha))
    c = None if cnts is None else np.ascontiguousarray(cnts, np.int32)
    lib().or_dirichlet_conj(seed, chain, sweep, tag, site0, alpha,
                            None if c is None else c.ctypes.data_as(C.c_void_p),
                            len(alpha), out)
    return out


def _sq(a):
    return np.ascontiguousarray(a, np.float64).copy()


def cholesky(a):
    a = _sq(a)
    lib().or_mat_cholesky(a, a.shape[0])
    return a


def lower_trmi(a):
    a = _sq(a)
    lib().or_mat_lower_trmi(a, a.shape[0])
    return a


def cholesky_invert(a):
    a = _sq(a)
    lib().or_mat_cholesky_invert(a, a.shape[0])
    return a


def trtrmm(l1, l2):
    l2 = _sq(l2)
    lib().or_mat_trtrmm(_sq(l1), l2, l2.shape[0])
    return l2


def trvm(l, v):
    v = _sq(v)
    lib().or_mat_trvm(_sq(l), v, len(v))
    return v


def mvnormal_ll(pt, mean, cov):
    return lib().or_mvnormal_ll(_sq(pt), _sq(mean), _sq(cov), len(pt))


def multi_lgamma(p, x):
    return lib().or_multi_lgamma(int(p), float(x))


def dirichlet_ll(pt, alpha):
    return lib().or_dirichlet_ll(_sq(pt), _sq(alpha), len(alpha))


def invwishart_ll(pt, df, scale):
    return lib().or_invwishart_ll(_sq(pt), int(df), _sq(scale), len(scale))


def mvnormal_conj(seed, chain, sweep, site0, mu0, cov0, s, cnt, cov):
    out = np.zeros(len(mu0))
    lib().or_mvnormal_mvnormal_conj(seed, chain, sweep, site0, _sq(mu0), _sq(cov0), _sq(s),
                                    int(cnt), _sq(cov), len(mu0), out)
    return out


def invwishart(seed, chain, sweep, site0, df, scale):
    out = np.zeros_like(_sq(scale))
    lib().or_invwishart_sample(seed, chain, sweep, site0, int(df), _sq(scale),
                               scale.shape[0], out)
    return out


def lda_counts(n_docs, k, v, doc_off, w, z):
    ct = np.zeros(n_docs * k, np.int32)
    cp = np.zeros(k * v, np.int32)
    lib().or_lda_counts(n_docs, k, v, np.ascontiguousarray(doc_off, np.int32),
                        np.ascontiguousarray(w, np.int32), np.ascontiguousarray(z, np.int32),
                        ct, cp)
    return ct.reshape(n_docs, k), cp.reshape(k, v)


def lda_zstep_f32(doc_off, w, theta32, phi_t32, seed, chain, sweep):
    n_docs, k = theta32.shape
    z = np.zeros(len(w), np.int32)
    lib().or_lda_zstep_f32(n_docs, k, np.ascontiguousarray(doc_off, np.int32),
                           np.ascontiguousarray(w, np.int32),
                           np.ascontiguousarray(theta32, np.float32),
                           np.ascontiguousarray(phi_t32, np.float32), seed, chain, sweep, z)
    return z


def lda_sweep(doc_off, w, alpha, beta, seed, chain, sweep, theta, phi, z):
    """In-place sweep on (theta[D,K], phi[K,V], z[T]) float64/int32 arrays."""
    n_docs, k = theta.shape
    v = phi.shape[1]
    lib().or_lda_sweep(n_docs, k, v, np.ascontiguousarray(doc_off, np.int32),
                       np.ascontiguousarray(w, np.int32), _sq(alpha), _sq(beta), seed, chain,
                       sweep, theta, phi, z)


class Hlr:
    """Keeps the numpy buffers alive for the C struct."""

    def __init__(self, x, y, prior_kind=0, prior_a=1.0, prior_b=1.0, has_intercept=False,
                 ref_compat=True):
        self.x = np.ascontiguousarray(x, np.float64)
        self.y = np.ascontiguousarray(y, np.int32)
        n, d = self.x.shape
        self.m = HlrModel(n, d, self.x.ctypes.data, self.y.ctypes.data, prior_kind,
                          prior_a, prior_b, int(has_intercept), int(ref_compat))
        self.d = d

    def data_ll(self, state, stable=True):
        return lib().or_hlr_data_ll(C.byref(self.m), _sq(state), int(stable))

    def data_grad(self, state, ref_form=False):
        g = np.zeros(self.d)
        gb = C.c_double(0.0)
        lib().or_hlr_data_grad(C.byref(self.m), _sq(state), g, C.byref(gb), int(ref_form))
        return g, gb.value

    def block_ll(self, state, block):
        return lib().or_hlr_block_ll(C.byref(self.m), _sq(state), block)

    def block_grad(self, state, block):
        dim = {0: 1, 1: 1, 2: self.d, 3: self.d + 1}[block]
        g = np.zeros(dim)
        lib().or_hlr_block_grad(C.byref(self.m), _sq(state), block, g)
        return g

    def hmc_step(self, state, block, block_id, sim_len, eps, fixed_steps, seed, chain, sweep):
        """In place on state ([theta, b, v], float64); returns info[7]."""
        info = np.zeros(7)
        lib().or_hmc_step(C.byref(self.m), state, block, block_id, sim_len, eps,
                          int(fixed_steps), seed, chain, sweep, info)
        return info

    def mwg_v_step(self, state, block_id, var, seed, chain, sweep):
        """Metropolis-within-Gibbs on v in place; returns info[4]."""
        info = np.zeros(4)
        lib().or_mwg_v_step(C.byref(self.m), state, block_id, var, seed, chain, sweep, info)
        return info

    def mwg_b_step(self, state, block_id, var, seed, chain, sweep):
        info = np.zeros(4)
        lib().or_mwg_b_step(C.byref(self.m), state, block_id, var, seed, chain, sweep, info)
        return info


class Gmm:
    """mu_kernel: 0 ConjGibbs, 1 ESlice, 2 HMC, 3 MWG (proposal MvNormal(mu[k],
    mu_prop_cov)), 4 RSlice."""

    def __init__(self, y, alpha, mu0, cov0, df, scale0, sample_pi=True, sample_cov=True,
                 mu_kernel=0, mu_sim_len=1.0, mu_step_size=0.05, mu_fixed_steps=0,
                 mu_prop_cov=None, ref_compat=True):
        self.y = np.ascontiguousarray(y, np.float64)
        n, d = self.y.shape
        k = len(alpha)
        self.alpha = _sq(alpha)
        self.mu0 = _sq(mu0)
        self.cov0 = _sq(cov0)
        self.scale0 = _sq(scale0)
        self.prop_cov = _sq(mu_prop_cov if mu_prop_cov is not None else cov0)
        self.m = GmmModel(n, d, k, self.y.ctypes.data, self.alpha.ctypes.data,
                          self.mu0.ctypes.data, self.cov0.ctypes.data, int(df),
                          self.scale0.ctypes.data, int(sample_pi), int(sample_cov),
                          int(mu_kernel), float(mu_sim_len), float(mu_step_size),
                          int(mu_fixed_steps), self.prop_cov.ctypes.data, int(ref_compat))

    def mu_move(self, mu, cov, z, seed, chain, sweep):
        """The model's mu kernel applied to mu (k x d, in place)."""
        return lib().or_gmm_mu_move(C.byref(self.m), mu, _sq(cov), np.ascontiguousarray(z, np.int32),
                                    seed, chain, sweep)

    def mu_fc(self, mu, cov, z):
        return lib().or_gmm_mu_fc(C.byref(self.m), np.ascontiguousarray(z, np.int32), _sq(mu), _sq(cov))

    def mu_grad(self, mu, cov, z):
        g = np.zeros_like(_sq(mu))
        lib().or_gmm_mu_grad(C.byref(self.m), np.ascontiguousarray(z, np.int32), _sq(mu), _sq(cov), g)
        return g

    def sweep(self, pis, mu, cov, z, seed, chain, sweep):
        lib().or_gmm_sweep(C.byref(self.m), pis, mu, cov, z, seed, chain, sweep)

    def zstep(self, pis, mu, cov, seed, chain, sweep):
        z = np.zeros(self.m.n, np.int32)
        lib().or_gmm_zstep(C.byref(self.m), _sq(pis), _sq(mu), _sq(cov), z, seed, chain, sweep)
        return z

    def ll_mod(self, pis, mu, cov, z):
        """The model's log density at (pis, mu, cov, z) (augur_get_curr_ll)."""
        return lib().or_gmm_ll_mod(C.byref(self.m), _sq(pis), _sq(mu), _sq(cov),
                                   np.ascontiguousarray(z, np.int32))
