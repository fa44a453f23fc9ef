"""augur_op_hlr_ll_grad_noalloc, the data term without allocation, against
augur_op_hlr_ll_grad (the entry point that allocates) on the same inputs:
two calls in a row, the second on a smaller input, so that whatever the
first call left in the shared scratch cannot leak into the second.

The caller's device buffers come from hipMalloc of the HIP runtime the
library itself links, called here through ctypes."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip():
    from augurv2_amd import _lib
    _lib.require_gpu()
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipDeviceSynchronize.argtypes = []
    return _lib, rt


class _Dev:
    """Device copies of host arrays (and output buffers) for one call."""

    def __init__(self, rt):
        self.rt, self.ptrs = rt, []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        self.ptrs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.rt.hipFree(p)


def _case(rng, n, d, chains):
    x = rng.standard_normal((n, d)).astype(np.float32)
    y = (rng.random(n) < 0.5).astype(np.int32)
    beta = rng.standard_normal((chains, d + 1)) * 0.3
    return x, y, beta


def test_noalloc_matches_the_allocating_entry_point(hip):
    _lib, rt = hip
    rng = np.random.default_rng(17)
    # second call smaller in rows, features and chains
    for n, d, chains, has_b in ((20000, 64, 5, 1), (777, 16, 3, 0)):
        x, y, beta = _case(rng, n, d, chains)
        ll_ref, g_ref = np.zeros(chains), np.zeros((chains, d + 1))
        _lib.call("augur_op_hlr_ll_grad", n, d, chains, _lib.ptr(x), _lib.ptr(y), _lib.ptr(beta), has_b,
                  _lib.ptr(ll_ref), _lib.ptr(g_ref))
        dev = _Dev(rt)
        try:
            px, py, pb = dev.put(x), dev.put(y), dev.put(beta)
            pll, pg = dev.put(np.full(chains, np.nan)), dev.put(np.full((chains, d + 1), np.nan))
            _lib.call("augur_op_hlr_ll_grad_noalloc", n, d, chains, px, py, pb, has_b, pll, pg, None)
            assert rt.hipDeviceSynchronize() == 0
            ll, g = dev.get(pll, ll_ref), dev.get(pg, g_ref)
        finally:
            dev.free()
        # the bar of the FP32 data term: log-density within 1e-5 relative,
        # gradient within 1e-5 of sum_n |a_n x_nj| per component
        np.testing.assert_allclose(ll, ll_ref, rtol=1e-5)
        x64 = np.concatenate([x.astype(np.float64), np.ones((n, 1))], axis=1)
        for c in range(chains):
            b = beta[c].copy()
            if not has_b:
                b[d] = 0.0
            a = y - 1.0 / (1.0 + np.exp(-(x64 @ b)))
            scale = np.abs(a) @ np.abs(x64)
            assert np.all(np.abs(g[c] - g_ref[c]) <= 1e-5 * scale + 1e-9), (n, d, c)


def test_noalloc_rejects_bad_arguments(hip):
    _lib, rt = hip
    with pytest.raises(_lib.AugurError):
        _lib.call("augur_op_hlr_ll_grad_noalloc", 100, 300, 1, None, None, None, 0, None, None, None)
    with pytest.raises(_lib.AugurError):
        _lib.call("augur_op_hlr_ll_grad_noalloc", 100, 8, 17, None, None, None, 0, None, None, None)
