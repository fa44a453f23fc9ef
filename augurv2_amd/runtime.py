"""Python handles over the C-ABI samplers (one object per augur_sampler*)."""
import ctypes as C

import numpy as np

from . import _lib as L


class Sampler:
    """Common operations of an augur_sampler handle."""

    family = None

    def __init__(self, handle, chains, shapes):
        self._h = handle
        self.chains = chains
        self._shapes = shapes  # name -> (shape per chain, dtype)

    # -- lifecycle --
    def close(self):
        if self._h:
            L.call("augur_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- parameters --
    def _spec(self, name):
        if name not in self._shapes:
            raise KeyError(f"{self.family} sampler has no parameter {name!r}")
        return self._shapes[name]

    def get(self, name, chain=0):
        """augur_cpy: current value of `name` for one chain."""
        shape, dt = self._spec(name)
        out = np.empty(shape, dtype=dt)
        L.call("augur_cpy", self._h, name.encode(), chain, L.ptr(out), out.nbytes)
        return out

    def set(self, name, value, chain=-1):
        """augur_set_pt: write `name` for one chain (-1: every chain)."""
        shape, dt = self._spec(name)
        a = np.ascontiguousarray(value, dtype=dt).reshape(shape)
        L.call("augur_set_pt", self._h, name.encode(), chain, L.ptr(a), a.nbytes)

    def step(self, n=1):
        """augur_step: run n sweeps on every chain."""
        L.call("augur_step", self._h, int(n))

    def log_density(self):
        out = np.zeros(self.chains)
        L.call("augur_get_curr_ll", self._h, L.ptr(out))
        return out

    def trace(self, name, capacity):
        L.call("augur_trace", self._h, name.encode(), int(capacity))

    def read_trace(self, name):
        shape, dt = self._spec(name)
        n = C.c_int(0)
        L.call("augur_trace_read", self._h, name.encode(), None, 0, C.byref(n))
        out = np.empty((n.value, self.chains) + tuple(shape), dtype=dt)
        L.call("augur_trace_read", self._h, name.encode(), L.ptr(out), out.nbytes, C.byref(n))
        return out

    def profile(self, period):
        """Event-time the dominant kernel every `period`-th sweep (0: off)."""
        L.call("augur_profile", self._h, int(period))

    def profile_read(self):
        """(average microseconds per timed launch, number of launches)."""
        avg = C.c_double(0.0)
        n = C.c_int(0)
        L.call("augur_profile_read", self._h, C.byref(avg), C.byref(n))
        return avg.value, n.value

    def sweeps(self):
        v = C.c_uint32(0)
        L.call("augur_sweep_count", self._h, C.byref(v))
        return v.value


class LdaSampler(Sampler):
    family = "lda"

    def __init__(self, n_topics, doc_off, w, alpha, beta, chains=1, seed=0):
        L.require_gpu()
        self.doc_off = L.i32(doc_off)
        self.w = L.i32(w)
        self.alpha = L.f64(alpha)
        self.beta = L.f64(beta)
        D = len(self.doc_off) - 1
        V = len(self.beta)
        d = L.LdaDesc(n_topics, D, V, chains, seed & 0xFFFFFFFF, L.ptr(self.doc_off).value,
                      L.ptr(self.w).value, L.ptr(self.alpha).value, L.ptr(self.beta).value)
        h = C.c_void_p()
        L.call("augur_lda_create", C.byref(d), C.byref(h))
        T = int(self.doc_off[-1])
        super().__init__(h, chains, {"theta": ((D, n_topics), np.float64),
                                     "phi": ((n_topics, V), np.float64),
                                     "z": ((T,), np.int32)})


    def predictive_ll(self, doc_off_ho, w_ho):
        """Per chain: sum over the held-out tokens (CSR over the model's
        documents) of log sum_k theta[d, k] phi[k, w] (FP64, on the device)."""
        off, w = L.i32(doc_off_ho), L.i32(w_ho)
        out = np.zeros(self.chains)
        L.call("augur_lda_predictive_ll", self._h, len(off) - 1, L.ptr(off), L.ptr(w), L.ptr(out))
        return out


class HlrSampler(Sampler):
    family = "hlr"

    def __init__(self, x, y, blocks, chains=1, seed=0, prior_kind=0, prior_a=1.0, prior_b=1.0,
                 has_intercept=False, ref_compat=True):
        """blocks: list of (block_kind, sim_len, step_size, fixed_steps)."""
        L.require_gpu()
        self.x = L.f32(x)
        self.y = L.i32(y)
        n, dim = self.x.shape
        desc = L.HlrDesc()
        desc.n_rows, desc.n_features, desc.chains, desc.seed = n, dim, chains, seed & 0xFFFFFFFF
        desc.x, desc.y = L.ptr(self.x).value, L.ptr(self.y).value
        desc.prior_kind, desc.prior_a, desc.prior_b = prior_kind, prior_a, prior_b
        desc.has_intercept, desc.ref_compat = int(has_intercept), int(ref_compat)
        if not 1 <= len(blocks) <= 4:
            raise ValueError("1..4 HMC blocks")
        desc.n_blocks = len(blocks)
        for i, (kind, sim_len, eps, fixed) in enumerate(blocks):
            desc.blocks[i] = L.HmcBlock(kind, sim_len, eps, int(fixed))
        h = C.c_void_p()
        L.call("augur_hlr_create", C.byref(desc), C.byref(h))
        self.n_blocks = len(blocks)
        shapes = {"state": ((dim + 2,), np.float64), "theta": ((dim,), np.float64),
                  "v": ((1,), np.float64),
                  # read-only: sigmoid(x[n].theta + b) per row, computed when read
                  "bias": ((n,), np.float64)}
        if has_intercept:
            shapes["b"] = ((1,), np.float64)
        super().__init__(h, chains, shapes)

    def enable_surrogate(self, block_index, hidden=256, train_sweeps=2000, max_pairs=500_000,
                         epochs=20, batch=512, lr=1e-3, steps=0, step_size=0.0,
                         min_accept=0.6, window=50, seed=0, linear_only=False):
        """Surrogate-gradient HMC for block `block_index` (an HMC [theta] block).
        steps = 0: surrogate trajectories follow the block's own (jittered)
        schedule; steps > 0: that many leapfrog steps of step_size.
        linear_only: the least-squares linear term without the network."""
        cfg = L.SurrogateCfg(block_index, hidden, train_sweeps, max_pairs, epochs, batch, lr, steps,
                             step_size, min_accept, window, seed & 0xFFFFFFFF, int(linear_only))
        L.call("augur_hlr_surrogate", self._h, C.byref(cfg))

    def surrogate_info(self):
        out = np.zeros(9)
        L.call("augur_hlr_surrogate_info", self._h, L.ptr(out), out.nbytes)
        keys = ("phase", "pairs", "loss", "train_ms", "fallbacks", "surrogate_sweeps",
                "window_accept", "collected_sweeps", "resid_frac")
        return dict(zip(keys, out.tolist()))

    def acceptance(self):
        """[block][chain] accepted / proposed."""
        out = np.zeros((self.n_blocks, self.chains, 2))
        L.call("augur_stats", self._h, L.ptr(out), out.nbytes)
        with np.errstate(invalid="ignore"):
            return out[..., 0] / out[..., 1]


class BhlrSampler(Sampler):
    """Many independent HLR datasets sampled together (BF16 X on the device)."""

    family = "bhlr"

    def __init__(self, xs, ys, chains=4, seed=0, v_block=(1.0, 0.05, 0), theta_block=(0.1, 0.005, 20),
                 prior_kind=0, prior_a=1.0, prior_b=1.0, ref_compat=False, rows_per_block=0,
                 use_graph=True):
        L.require_gpu()
        self.xs = [L.f32(x) for x in xs]
        self.ys = [L.i32(y).ravel() for y in ys]
        n = np.array([x.shape[0] for x in self.xs], np.int32)
        d = np.array([x.shape[1] for x in self.xs], np.int32)
        self.n_rows, self.dims = n, d
        xp = (C.c_void_p * len(xs))(*[L.ptr(x).value for x in self.xs])
        yp = (C.c_void_p * len(ys))(*[L.ptr(y).value for y in self.ys])
        self._keep = (xp, yp)
        desc = L.BhlrDesc()
        desc.n_datasets = len(xs)
        desc.n_rows, desc.n_features = L.ptr(n).value, L.ptr(d).value
        desc.x, desc.y = C.cast(xp, C.c_void_p).value, C.cast(yp, C.c_void_p).value
        desc.chains, desc.seed = chains, seed & 0xFFFFFFFF
        desc.prior_kind, desc.prior_a, desc.prior_b = prior_kind, prior_a, prior_b
        desc.ref_compat = int(ref_compat)
        # v_block: (sim_len, eps, fixed_steps) for HMC, or ("mwg", variance)
        if v_block[0] == "mwg":
            desc.v_block = L.HmcBlock(L.MWG_BLOCK_V, 0.0, float(v_block[1]), 0)
        else:
            desc.v_block = L.HmcBlock(L.HMC_BLOCK_V, *v_block[:2], int(v_block[2]))
        desc.theta_block = L.HmcBlock(L.HMC_BLOCK_THETA, *theta_block[:2], int(theta_block[2]))
        desc.rows_per_block, desc.use_graph = int(rows_per_block), int(use_graph)
        h = C.c_void_p()
        L.call("augur_bhlr_create", C.byref(desc), C.byref(h))
        self.offsets = np.concatenate([[0], np.cumsum(d + 2)])[:-1]
        shapes = {"state": ((int(np.sum(d + 2)),), np.float64)}
        for i in range(len(xs)):
            shapes[f"state:{i}"] = ((int(d[i]) + 2,), np.float64)
        super().__init__(h, chains, shapes)

    def adapt(self, n_sweeps, target_accept=0.8):
        """Warmup sweeps that adapt each dataset's theta step size (and the
        MWG proposal variance of v), then freeze them."""
        L.call("augur_bhlr_adapt", self._h, int(n_sweeps), float(target_accept))

    def step_sizes(self):
        """(theta step size, v proposal variance) per dataset."""
        k = len(self.xs)
        eps, vv = np.zeros(k), np.zeros(k)
        L.call("augur_bhlr_step_sizes", self._h, L.ptr(eps), L.ptr(vv), k)
        return eps, vv

    def tuning(self):
        rows = C.c_int(0)
        n = C.c_int(0)
        cr = np.zeros(16, np.int32)
        cu = np.zeros(16, np.float32)
        L.call("augur_bhlr_tuning", self._h, C.byref(rows), C.byref(n), L.ptr(cr), L.ptr(cu), 16)
        return rows.value, dict(zip(cr[:n.value].tolist(), cu[:n.value].tolist()))

    def surrogate(self, train_sweeps, ridge=1e-6, min_accept=0.3, window=50):
        """Linear gradient surrogate for every dataset: the next train_sweeps
        sweeps collect (theta, data gradient) pairs, each dataset's map is
        then fitted, and later sweeps run surrogate trajectories with an
        exact Metropolis test; a dataset below min_accept over a window of
        sweeps returns to exact HMC."""
        cfg = L.BsurrCfg(int(train_sweeps), float(ridge), float(min_accept), int(window))
        L.call("augur_bhlr_surrogate", self._h, C.byref(cfg))

    def surrogate_info(self):
        info = np.zeros(7)
        L.call("augur_bhlr_surrogate_info", self._h, L.ptr(info), info.nbytes)
        keys = ("phase", "datasets_on_surrogate", "fallbacks", "fit_ms", "pairs_per_dataset",
                "surrogate_sweeps", "collection_sweeps")
        return dict(zip(keys, info.tolist()))

    def surrogate_modes(self):
        m = np.zeros(len(self.xs), np.int32)
        L.call("augur_bhlr_surrogate_modes", self._h, L.ptr(m), len(self.xs))
        return m

    def surrogate_set_mode(self, i, mode):
        """Dataset i on exact HMC (0) or on its fitted surrogate (1)."""
        L.call("augur_bhlr_surrogate_set_mode", self._h, int(i), int(mode))

    def surrogate_map(self, i):
        """(W, m, gbar) of dataset i: g(theta) ~ gbar + W (theta - m)."""
        d = int(self.dims[i])
        w, m, g = np.zeros((d, d)), np.zeros(d), np.zeros(d)
        L.call("augur_bhlr_surrogate_map", self._h, int(i), L.ptr(w), L.ptr(m), L.ptr(g))
        return w, m, g

    def surrogate_pairs(self, i):
        """Dataset i's fitting pairs (theta, data gradient), FP32 rows in
        (slot, chain) order."""
        d = int(self.dims[i])
        n = C.c_int(0)
        L.call("augur_bhlr_surrogate_pairs", self._h, int(i), None, None, 0, C.byref(n))
        t = np.zeros((n.value, d), np.float32)
        g = np.zeros((n.value, d), np.float32)
        L.call("augur_bhlr_surrogate_pairs", self._h, int(i), L.ptr(t), L.ptr(g), n.value, C.byref(n))
        return t, g

    def acceptance_counts(self):
        """[block (v, theta)][chain][dataset][accepted, proposed] since creation."""
        k = len(self.xs)
        out = np.zeros((2, self.chains, k, 2))
        L.call("augur_stats", self._h, L.ptr(out), out.nbytes)
        return out

    def acceptance(self):
        """[block (v, theta)][chain][dataset] acceptance rates."""
        out = self.acceptance_counts()
        with np.errstate(invalid="ignore"):
            return out[..., 0] / out[..., 1]


class GmmSampler(Sampler):
    family = "gmm"

    def __init__(self, y, alpha, mu0, cov0, df, scale0, chains=1, seed=0, sample_pi=True,
                 sample_cov=True, mu_kernel=0, mu_sim_len=1.0, mu_step_size=0.05,
                 mu_fixed_steps=0, mu_prop_cov=None, ref_compat=True):
        """mu_kernel: L.GMM_MU_* (ConjGibbs, ESlice, HMC, MWG, RSlice)."""
        L.require_gpu()
        self.y = L.f32(y)
        n, dim = self.y.shape
        self.alpha, self.mu0 = L.f64(alpha), L.f64(mu0)
        self.cov0, self.scale0 = L.f64(cov0), L.f64(scale0)
        self.prop_cov = L.f64(mu_prop_cov) if mu_prop_cov is not None else None
        k = len(self.alpha)
        desc = L.GmmDesc(n, dim, k, chains, seed & 0xFFFFFFFF, L.ptr(self.y).value,
                         L.ptr(self.alpha).value, L.ptr(self.mu0).value, L.ptr(self.cov0).value,
                         int(df), L.ptr(self.scale0).value, int(sample_pi), int(sample_cov),
                         int(mu_kernel), float(mu_sim_len), float(mu_step_size),
                         int(mu_fixed_steps),
                         L.ptr(self.prop_cov).value if self.prop_cov is not None else None,
                         int(ref_compat))
        h = C.c_void_p()
        L.call("augur_gmm_create", C.byref(desc), C.byref(h))
        super().__init__(h, chains, {"pis": ((k,), np.float64), "mu": ((k, dim), np.float64),
                                     "cov": ((k, dim, dim), np.float64), "z": ((n,), np.int32)})

    def predictive_ll(self, y_test):
        """Per chain: sum over test points of log sum_k pis_k N(y | mu_k, cov_k)
        (FP64, on the device)."""
        yt = L.f64(y_test)
        out = np.zeros(self.chains)
        L.call("augur_gmm_predictive_ll", self._h, yt.shape[0], L.ptr(yt), L.ptr(out))
        return out

    def mu_acceptance(self):
        """Per chain: accepted / proposed moves of a non-conjugate mu kernel
        (ESlice counts a move when the slice returned a new point)."""
        out = np.zeros(1 + 2 * self.chains)
        L.call("augur_stats", self._h, L.ptr(out), out.nbytes)
        a = out[1:].reshape(self.chains, 2)
        with np.errstate(invalid="ignore"):
            return a[:, 0] / a[:, 1]


class GenSampler(Sampler):
    """A generic model: the passes generated by pyaugur.lower.Lowering,
    compiled with hipRTC, under HMC and DiscGibbs blocks."""

    family = "gen"

    def __init__(self, lowering, blocks, chains=1, seed=0):
        """blocks: list of (parameter names, sim_len, step_size, fixed_steps)
        for HMC, or ("DiscGibbs", name) for a discrete parameter."""
        L.require_gpu()
        lw = lowering
        self.P = lw.P
        self.NZ = getattr(lw, "NZ", 1)
        self.layout = dict(lw.layout)
        self.zlayout = dict(getattr(lw, "zlayout", {}))
        self._src = lw.source.encode()
        self._kernel = lw.kernel_name.encode()
        self._reals, self._ints = L.f64(lw.reals), L.i32(lw.ints)
        names = [n for n in lw.param_names if n in self.layout]
        znames = [n for n in lw.param_names if n in self.zlayout]
        self._names = [n.encode() for n in names]
        self._znames = [n.encode() for n in znames]
        name_arr = (C.c_char_p * max(1, len(names)))(*self._names)
        zname_arr = (C.c_char_p * max(1, len(znames)))(*self._znames)

        def _size(shape):
            return int(np.prod(shape)) if shape else 1
        offs = L.i32([self.layout[n][0] for n in names] or [0])
        sizes = L.i32([_size(self.layout[n][1]) for n in names] or [1])
        zoffs = L.i32([self.zlayout[n][0] for n in znames] or [0])
        zsizes = L.i32([_size(self.zlayout[n][1]) for n in znames] or [1])
        nb = max(1, len(blocks))
        masks = np.zeros((nb, self.P), np.uint8)
        sim_len, eps = np.ones(nb), np.ones(nb)
        fixed, kinds = np.zeros(nb, np.int32), np.zeros(nb, np.int32)
        for b, blk in enumerate(blocks):
            if blk[0] == "DiscGibbs":
                kinds[b] = L.GEN_DISC
                continue
            vars_, sl, ep, fx = blk
            for v in vars_:
                off, shape = self.layout[v]
                masks[b, off:off + _size(shape)] = 1
            sim_len[b], eps[b], fixed[b] = sl, ep, fx
        # the blocks' own kernels, when the lowering generated them for
        # exactly these blocks
        bk = getattr(lw, "block_kernels", [])
        if blocks and len(bk) == len(blocks):
            kn = [k[0].encode() for k in bk]
            kernel_arr = (C.c_char_p * len(kn))(*kn)
            grids = L.i32([k[1] for k in bk])
            kptr, gptr = C.cast(kernel_arr, C.c_void_p).value, L.ptr(grids).value
        elif any(k == L.GEN_DISC for k in kinds[:len(blocks)]):
            raise ValueError("DiscGibbs blocks need the lowering's kernels for the same blocks")
        else:
            kn, kernel_arr, grids, kptr, gptr = [], None, None, None, None
        self._keep = (name_arr, zname_arr, offs, sizes, zoffs, zsizes, masks, sim_len, eps, fixed, kinds,
                      kn, kernel_arr, grids)
        desc = L.GenDesc(self._src, self._kernel, self.P, lw.grid, chains, seed & 0xFFFFFFFF,
                         L.ptr(self._reals).value, self._reals.size, L.ptr(self._ints).value,
                         self._ints.size, len(names), C.cast(name_arr, C.c_void_p).value,
                         L.ptr(offs).value, L.ptr(sizes).value, len(blocks), L.ptr(masks).value,
                         L.ptr(sim_len).value, L.ptr(eps).value, L.ptr(fixed).value, kptr, gptr,
                         L.ptr(kinds).value, self.NZ, len(znames),
                         C.cast(zname_arr, C.c_void_p).value, L.ptr(zoffs).value, L.ptr(zsizes).value)
        h = C.c_void_p()
        L.call("augur_gen_create", C.byref(desc), C.byref(h))
        self.n_blocks = len(blocks)
        self.kinds = kinds[:len(blocks)].tolist()
        shapes = {"state": ((self.P,), np.float64), "zstate": ((self.NZ,), np.int32)}
        for n in names:
            shape = self.layout[n][1]
            shapes[n] = (shape if shape else (1,), np.float64)
        for n in znames:
            shape = self.zlayout[n][1]
            shapes[n] = (shape if shape else (1,), np.int32)
        super().__init__(h, chains, shapes)

    def eval(self, q):
        """(log joint per chain, gradient chains x P) at the points q (chains x
        P) and the current discrete values."""
        qa = L.f64(q).reshape(self.chains, self.P)
        ll = np.zeros(self.chains)
        g = np.zeros((self.chains, self.P))
        L.call("augur_gen_eval", self._h, L.ptr(qa), L.ptr(ll), L.ptr(g))
        return ll, g

    def acceptance(self):
        """[block][chain] accepted / proposed (HMC blocks; NaN for DiscGibbs)."""
        out = np.zeros((self.n_blocks, self.chains, 2))
        L.call("augur_stats", self._h, L.ptr(out), out.nbytes)
        with np.errstate(invalid="ignore"):
            return out[..., 0] / out[..., 1]
