"""ctypes binding of libaugur_hip.so (declared in include/augur_hip.h).

The library is loaded from the package directory; a missing library or a
missing GPU raises -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaugur_hip.so")

AUGUR_OK = 0
STATUS_NAMES = {1: "invalid argument", 2: "HIP error", 3: "invalid state", 4: "unsupported"}

HMC_BLOCK_V, HMC_BLOCK_B, HMC_BLOCK_THETA, HMC_BLOCK_B_THETA = 0, 1, 2, 3
MWG_BLOCK_V = 4  # proposal Normal(v, step_size): step_size is the variance
MWG_BLOCK_B = 5  # proposal Normal(b, step_size) on the intercept


class AugurError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class LdaDesc(C.Structure):
    _fields_ = [("n_topics", C.c_int), ("n_docs", C.c_int), ("n_vocab", C.c_int),
                ("chains", C.c_int), ("seed", C.c_uint32), ("doc_off", C.c_void_p),
                ("w", C.c_void_p), ("alpha", C.c_void_p), ("beta", C.c_void_p)]


class HmcBlock(C.Structure):
    _fields_ = [("block", C.c_int), ("sim_len", C.c_double), ("step_size", C.c_double),
                ("fixed_steps", C.c_int)]


class HlrDesc(C.Structure):
    _fields_ = [("n_rows", C.c_int), ("n_features", C.c_int), ("chains", C.c_int),
                ("seed", C.c_uint32), ("x", C.c_void_p), ("y", C.c_void_p),
                ("prior_kind", C.c_int), ("prior_a", C.c_double), ("prior_b", C.c_double),
                ("has_intercept", C.c_int), ("ref_compat", C.c_int), ("n_blocks", C.c_int),
                ("blocks", HmcBlock * 4)]


class BhlrDesc(C.Structure):
    _fields_ = [("n_datasets", C.c_int), ("n_rows", C.c_void_p), ("n_features", C.c_void_p),
                ("x", C.c_void_p), ("y", C.c_void_p), ("chains", C.c_int), ("seed", C.c_uint32),
                ("prior_kind", C.c_int), ("prior_a", C.c_double), ("prior_b", C.c_double),
                ("ref_compat", C.c_int), ("v_block", HmcBlock), ("theta_block", HmcBlock),
                ("rows_per_block", C.c_int), ("use_graph", C.c_int)]


class SurrogateCfg(C.Structure):
    _fields_ = [("block_index", C.c_int), ("hidden", C.c_int), ("train_sweeps", C.c_int),
                ("max_pairs", C.c_int), ("epochs", C.c_int), ("batch", C.c_int),
                ("lr", C.c_double), ("steps", C.c_int), ("step_size", C.c_double),
                ("min_accept", C.c_double), ("window", C.c_int), ("seed", C.c_uint32),
                ("linear_only", C.c_int)]


class BsurrCfg(C.Structure):
    _fields_ = [("train_sweeps", C.c_int), ("ridge", C.c_double), ("min_accept", C.c_double),
                ("window", C.c_int)]


class GmmDesc(C.Structure):
    _fields_ = [("n_points", C.c_int), ("dim", C.c_int), ("n_components", C.c_int),
                ("chains", C.c_int), ("seed", C.c_uint32), ("y", C.c_void_p),
                ("alpha", C.c_void_p), ("mu0", C.c_void_p), ("cov0", C.c_void_p),
                ("df", C.c_int), ("scale0", C.c_void_p), ("sample_pi", C.c_int),
                ("sample_cov", C.c_int), ("mu_kernel", C.c_int), ("mu_sim_len", C.c_double),
                ("mu_step_size", C.c_double), ("mu_fixed_steps", C.c_int),
                ("mu_prop_cov", C.c_void_p), ("ref_compat", C.c_int)]


class GenDesc(C.Structure):
    _fields_ = [("source", C.c_char_p), ("kernel", C.c_char_p), ("n_params", C.c_int),
                ("grid", C.c_int), ("chains", C.c_int), ("seed", C.c_uint32),
                ("reals", C.c_void_p), ("n_reals", C.c_size_t), ("ints", C.c_void_p),
                ("n_ints", C.c_size_t), ("n_names", C.c_int), ("names", C.c_void_p),
                ("offsets", C.c_void_p), ("sizes", C.c_void_p), ("n_blocks", C.c_int),
                ("masks", C.c_void_p), ("sim_len", C.c_void_p), ("step_size", C.c_void_p),
                ("fixed_steps", C.c_void_p), ("block_kernels", C.c_void_p),
                ("block_grids", C.c_void_p), ("block_kinds", C.c_void_p), ("n_zints", C.c_int),
                ("n_znames", C.c_int), ("znames", C.c_void_p), ("zoffsets", C.c_void_p),
                ("zsizes", C.c_void_p)]


GEN_HMC, GEN_DISC = 0, 1


GMM_MU_CONJ, GMM_MU_ESLICE, GMM_MU_HMC, GMM_MU_MWG, GMM_MU_RSLICE, GMM_MU_NUTS = 0, 1, 2, 3, 4, 5


_H = C.c_void_p
_P = C.c_void_p
_U = C.c_uint32
_I = C.c_int
_S = C.c_size_t
_D = C.c_double

# name -> argtypes (all return augur_status unless listed in _RESTYPE)
SIGNATURES = {
    "augur_last_error": [],
    "augur_version": [],
    "augur_device_count": [_P],
    "augur_lda_create": [_P, _P],
    "augur_hlr_create": [_P, _P],
    "augur_gmm_create": [_P, _P],
    "augur_bhlr_create": [_P, _P],
    "augur_gen_create": [_P, _P],
    "augur_gen_eval": [_H, _P, _P, _P],
    "augur_hlr_surrogate": [_H, _P],
    "augur_hlr_surrogate_info": [_H, _P, _S],
    "augur_op_mlp_loss_grad": [_I, _I, _I, _P, _P, _P, _P, _P, _P],
    "augur_op_mlp_train": [_I, _I, _I, _P, _P, _I, _I, C.c_double, _U, _P, _P],
    "augur_bhlr_tuning": [_H, _P, _P, _P, _P, _I],
    "augur_bhlr_adapt": [_H, _I, _D],
    "augur_bhlr_step_sizes": [_H, _P, _P, _I],
    "augur_bhlr_surrogate": [_H, _P],
    "augur_bhlr_surrogate_info": [_H, _P, _S],
    "augur_bhlr_surrogate_modes": [_H, _P, _I],
    "augur_bhlr_surrogate_set_mode": [_H, _I, _I],
    "augur_bhlr_surrogate_map": [_H, _I, _P, _P, _P],
    "augur_bhlr_surrogate_pairs": [_H, _I, _P, _P, _I, _P],
    "augur_lda_predictive_ll": [_H, _I, _P, _P, _P],
    "augur_gmm_predictive_ll": [_H, _I, _P, _P],
    "augur_op_lda_predictive_ll": [_I, _I, _I, _P, _P, _P, _P, _P],
    "augur_op_gmm_predictive_ll": [_I, _I, _P, _P, _P, _I, _P, _P],
    "augur_set_pt": [_H, C.c_char_p, _I, _P, _S],
    "augur_cpy": [_H, C.c_char_p, _I, _P, _S],
    "augur_step": [_H, _I],
    "augur_get_curr_ll": [_H, _P],
    "augur_trace": [_H, C.c_char_p, _I],
    "augur_trace_read": [_H, C.c_char_p, _P, _S, _P],
    "augur_stats": [_H, _P, _S],
    "augur_sweep_count": [_H, _P],
    "augur_profile": [_H, _I],
    "augur_profile_read": [_H, _P, _P],
    "augur_destroy": [_H],
    "augur_op_philox": [_P, _P, _I, _P],
    "augur_op_stream_bandwidth": [_S, _I, _P, _P],
    "augur_op_gamma": [_U, _U, _U, _U, _P, _I, _P],
    "augur_op_dirichlet_conj": [_U, _U, _U, _U, _P, _P, _I, _I, _P],
    "augur_op_lda_zstep": [_I, _I, _I, _P, _P, _P, _P, _U, _U, _U, _P, _P, _P],
    "augur_op_hlr_ll_grad": [_I, _I, _I, _P, _P, _P, _I, _P, _P],
    "augur_op_hlr_ll64": [_I, _I, _I, _P, _P, _P, _I, _P],
    "augur_op_hlr_ll_grad_noalloc": [_I, _I, _I, _P, _P, _P, _I, _P, _P, _P],
    "augur_op_gmm_zstep": [_I, _I, _I, _P, _P, _P, _P, _U, _U, _U, _P],
    "augur_op_gmm_conj": [_I, _I, _P, _P, _P, _I, _P, _P, _P, _P, _I, _I, _U, _U, _U, _P, _P,
                          _P],
}
_RESTYPE = {"augur_last_error": C.c_char_p, "augur_version": C.c_char_p}

_lib = None


def load():
    """Load the HIP sampler library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                "g.build()'` (hipcc, gfx950)")
        lib = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, C.c_int)
        _lib = lib
    return _lib


def check(status):
    if status != AUGUR_OK:
        raise AugurError(status, load().augur_last_error().decode())


def call(name, *args):
    check(getattr(load(), name)(*args))


def device_count():
    n = C.c_int(0)
    load().augur_device_count(C.byref(n))
    return n.value


def require_gpu():
    if device_count() < 1:
        raise AugurError(2, "no HIP device visible: the augurv2_amd sampler runs on an MI355X GPU")


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)
