"""Generic lowering of continuous models to a HIP data pass.

This plays the role of the reference's code generators for models that no
hand-written sampler family covers.  The reference's generators are
compiler/augur/src/CudaC/CgCuda.hs:371-516 (map / map-reduce kernels), with
Core/CgLike.hs:90-267 (the log joint) and Core/CgGrad.hs:151-376 (its
reverse-mode gradient).  Here a model whose parameters are all continuous
becomes one HIP kernel source.  It is compiled at run time with hipRTC and
driven by the library's device-resident HMC (augur_gen_create).

The generated kernel ``augur_gen_pass`` runs on a grid of (G, chains)
blocks of 256 threads, G a multiple of 8 with about four points of the
longest grid per thread.  The block at dispatch position L takes row tile
(L / 8 / chains) * 8 + L % 8 of chain (L / 8) % chains: dispatch goes
round-robin over the 8 XCDs, so the chains of one tile run on one XCD and
its L2 serves their common data reads (measured on hlr.rv at N = 100k,
D = 64, 4 chains: 69 us per pass against 248 us with one row per thread
and blocks in chain-major order).  Each block does the following.
  - It loads its chain's parameter vector q (P reals) into LDS.
  - It walks every declaration ``x[i..] ~ Dist(args) for i <- 0 until n``
    as a grid-stride loop over the flattened grid.  A declaration without a
    grid runs on the chain's first thread only.
  - At each point it adds the log density and the reverse-mode
    contributions to d ll / dq, in FP64.  A contribution to a coordinate
    known when the kernel is generated goes to a register accumulator.  One
    whose coordinate depends on the grid index or on data (theta[k],
    alpha[g[n]]) is an LDS atomic add.
  - It sums the register accumulators over the block in a fixed order.
  - It writes one row of P + 1 partials (the gradient, then ll).
Besides the log joint (``augur_gen_pass``, for the log density and the
pass at given points), each HMC block b gets ``augur_gen_pass_b<b>`` over
only the declarations that mention its variables: its full conditional, so
that a block whose variables touch no data never reads the data.
The library sums the G rows of each chain in a fixed order.  The
result is deterministic whenever no coordinate receives two LDS atomic adds
from one block.

Layout
  - The parameters come in declaration order, each flattened row-major.
    Together they form q; chain c's vector is q[c * P .. c * P + P).
  - Real data and real hyper-parameter arrays are packed into one FP64
    buffer R.  Int arrays are packed into one int32 buffer I.  Arrays of
    rank 2 are stored column-major, so that x[n][k] at consecutive n are
    adjacent and a wave's loads coalesce.
  - Scalar hyper-parameters and every size are constants in the source.

Supported
  - Parameters: Normal, Exponential, Gamma, InvGamma, Beta and Uniform, of
    rank 0-2 over rectangular grids.
  - Data: those distributions plus Bernoulli and Poisson.
  - Expressions: + - * /, sigmoid, dotprod, literals, and projections
    indexed by grid variables, literals or int data.
  - Schedules: HMC blocks.
  - Anything else raises LoweringError.
"""
import math

import numpy as np

PARAM_DISTS = {"Normal", "Exponential", "Gamma", "InvGamma", "Beta", "Uniform"}
DISC_DISTS = {"Categorical", "Bernoulli"}   # discrete parameters (DiscGibbs)
DATA_DISTS = PARAM_DISTS | {"Bernoulli", "Poisson"}
INT_DISTS = {"Bernoulli", "Poisson", "Categorical"}
ARITY = {"Normal": 2, "Exponential": 1, "Gamma": 2, "InvGamma": 2, "Beta": 2, "Uniform": 2,
         "Bernoulli": 1, "Poisson": 1, "Categorical": 1}
MAX_SUPPORT = 64       # values of a discrete parameter
THREADS = 256
MAX_GRID = 512         # a multiple of 8
ROWS_PER_THREAD = 4
MAX_PARAMS = 3000      # q and the gradient row are held in LDS
MAX_REG_SLOTS = 96     # FP64 register accumulators per thread
UNROLL_CAP = 128       # longest dotprod whose adjoint loop is unrolled into registers


class LoweringError(ValueError):
    """The model is outside what the generic lowering supports."""


def _fail(msg):
    raise LoweringError(msg)


def _free_vars(e, out):
    """Names an expression mentions."""
    k = e[0]
    if k == "var":
        out.add(e[1])
    elif k == "proj":
        _free_vars(e[1], out)
        for i in e[2]:
            _free_vars(i, out)
    elif k in ("prim", "dist"):
        for a in e[2]:
            _free_vars(a, out)
    elif k == "bin":
        _free_vars(e[2], out)
        _free_vars(e[3], out)
    return out


def _lit(v):
    """C++ literal that parses to exactly the double v."""
    v = float(v)
    if math.isnan(v) or math.isinf(v):
        _fail(f"non-finite constant {v}")
    return repr(v) if ("e" in repr(v) or "." in repr(v)) else repr(v) + ".0"


# ---- the device prelude: densities and derivatives of cbits/augur_dist.c ----

PRELUDE = r"""
#define GEN_LARGE 1e8
// cbits/augur_dist.c:372-394 (kRefCompat: the reference's -log var and
// -0.5 var forms; otherwise the textbook -0.5 log var and -0.5 / var)
__device__ inline double gen_normal_ll(double pt, double mean, double var) {
  if (var > 0.0) { const double t = pt - mean; return -(kRefCompat ? 1.0 : 0.5) * log(var) - 0.5 * t * t / var; }
  return -GEN_LARGE;
}
__device__ inline double gen_normal_dotpt(double pt, double mean, double var) { return -(pt - mean) / var; }
__device__ inline double gen_normal_dotarg1(double pt, double mean, double var) { return (pt - mean) / var; }
__device__ inline double gen_normal_dotarg2(double pt, double mean, double var) {
  const double t = pt - mean;
  return (kRefCompat ? -0.5 * var : -0.5 / var) + 0.5 * (t * t) / (var * var);
}
// :59-71
__device__ inline double gen_bernoulli_ll(int pt, double bias) {
  if (bias >= 0.0 && bias <= 1.0) return log(pt * bias + (1 - pt) * (1.0 - bias));
  return -GEN_LARGE;
}
__device__ inline double gen_bernoulli_dotarg1(int pt, double bias) {
  return (double)(pt + pt - 1) / (pt * bias + (1 - pt) * (1.0 - bias));
}
// :124-139 (sum of log k for k = 1..pt as lgamma(pt + 1))
__device__ inline double gen_poisson_ll(int pt, double rate) { return pt * log(rate) - rate - lgamma(pt + 1.0); }
__device__ inline double gen_poisson_dotarg1(int pt, double rate) { return rate > 0.0 ? pt / rate - 1.0 : -GEN_LARGE; }
// :231-256
__device__ inline double gen_exponential_ll(double pt, double rate) {
  return (rate > 0.0 && pt > 0.0) ? log(rate) - rate * pt : -GEN_LARGE;
}
__device__ inline double gen_exponential_dotpt(double pt, double rate) { return (rate > 0.0 && pt > 0.0) ? -rate : 0.0; }
__device__ inline double gen_exponential_dotarg1(double pt, double rate) {
  return (rate > 0.0 && pt > 0.0) ? 1.0 / rate - pt : 0.0;
}
// digamma: recurrence psi(x) = psi(x + 1) - 1 / x up to x >= 6, then the
// asymptotic series ln x - 1/(2x) - sum B_2k / (2k x^2k) (Abramowitz and
// Stegun 6.3.5, 6.3.18), five terms
__device__ inline double gen_digamma(double x) {
  double r = 0.0;
  while (x < 6.0) { r -= 1.0 / x; x += 1.0; }
  const double f = 1.0 / (x * x);
  const double s = f * (1.0 / 12 - f * (1.0 / 120 - f * (1.0 / 252 - f * (1.0 / 240 - f * (1.0 / 132)))));
  return r + log(x) - 0.5 / x - s;
}
// :265-285 (the log of the pdf term by term)
__device__ inline double gen_gamma_ll(double pt, double shape, double rate) {
  return shape * log(rate) - lgamma(shape) + (shape - 1.0) * log(pt) - rate * pt;
}
__device__ inline double gen_gamma_dotpt(double pt, double shape, double rate) { return ((shape - 1.0) / pt) - rate; }
__device__ inline double gen_gamma_dotarg1(double pt, double shape, double rate) {
  return -gen_digamma(shape) + log(rate) + log(pt);
}
__device__ inline double gen_gamma_dotarg2(double pt, double shape, double rate) { return (shape / rate) - pt; }
// :345-359
__device__ inline double gen_invgamma_ll(double pt, double shape, double scale) {
  return shape * log(scale) - lgamma(shape) - (shape + 1.0) * log(pt) - scale / pt;
}
__device__ inline double gen_invgamma_dotpt(double pt, double shape, double scale) {
  return -(shape + 1.0) / pt + scale / (pt * pt);
}
__device__ inline double gen_invgamma_dotarg1(double pt, double shape, double scale) {
  return log(scale) - gen_digamma(shape) - log(pt);
}
__device__ inline double gen_invgamma_dotarg2(double pt, double shape, double scale) { return shape / scale - 1.0 / pt; }
// :158-177 (the log of the pdf term by term)
__device__ inline double gen_beta_ll(double pt, double a, double b) {
  return lgamma(a + b) - lgamma(a) - lgamma(b) + (a - 1.0) * log(pt) + (b - 1.0) * log(1.0 - pt);
}
__device__ inline double gen_beta_dotpt(double pt, double a, double b) { return (a - 1.0) / pt - (b - 1.0) / (1.0 - pt); }
__device__ inline double gen_beta_dotarg1(double pt, double a, double b) {
  return gen_digamma(a + b) - gen_digamma(a) + log(pt);
}
__device__ inline double gen_beta_dotarg2(double pt, double a, double b) {
  return gen_digamma(a + b) - gen_digamma(b) + log(1.0 - pt);
}
// :426-455
__device__ inline double gen_uniform_ll(double pt, double l, double u) { return (l < pt && pt < u) ? -log(u - l) : -GEN_LARGE; }
__device__ inline double gen_uniform_dotpt(double pt, double l, double u) { return 0.0; }
__device__ inline double gen_uniform_dotarg1(double pt, double l, double u) { return (l < pt && pt < u) ? 1.0 / (u - l) : 0.0; }
__device__ inline double gen_uniform_dotarg2(double pt, double l, double u) { return (l < pt && pt < u) ? -1.0 / (u - l) : 0.0; }

// Philox4x32-10 (Salmon et al., SC'11) with counter (site, sub, sweep, tag)
// and key (seed, chain), as the library's RNG; u = (w + 0.5) 2^-32
__device__ inline double gen_u01(unsigned seed, unsigned chain, unsigned site, unsigned sweep,
                                 unsigned tag) {
  unsigned c0 = site, c1 = 0u, c2 = sweep, c3 = tag, k0 = seed, k1 = chain;
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const unsigned long long p0 = (unsigned long long)0xD2511F53u * c0;
    const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c2;
    const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
    const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
    const unsigned n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
  }
  return ((double)c0 + 0.5) * 2.3283064365386963e-10;
}

__device__ inline double gen_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
"""


# ---- values of the code generator ----------------------------------------------

class _Val:
    """A generated value: a scalar (C++ expression ``code``) or a vector
    view (``elem(k)`` gives element k's expression, ``coord(k)`` its
    coordinate in q when the vector is a parameter's slice)."""

    def __init__(self, code=None, dep=False, op=None, kids=(), coord=None, length=None, elem=None,
                 vcoord=None):
        self.code = code
        self.dep = dep
        self.op = op
        self.kids = kids
        self.coord = coord        # scalar parameter element: int or C++ int expression
        self.length = length      # vector views
        self.elem = elem
        self.vcoord = vcoord

    @property
    def is_vec(self):
        return self.length is not None


class _Sym:
    """What a name denotes inside the model."""

    def __init__(self, kind, shape, is_int=False, value=None, off=None):
        self.kind = kind          # "hyper" | "param" | "dparam" (discrete) | "data"
        self.shape = shape        # () | (n,) | (n, m)
        self.is_int = is_int
        self.value = value        # hyper scalars: the constant
        self.off = off            # offset into q (params), Z (discrete) or R / I (arrays)
        self.support = 0          # discrete parameters: number of values


class Lowering:
    """Lowers one model (parsed by rvparse) bound to its hyper-parameters and
    data.  Attributes after construction:
      source, kernel_name, P, grid, reals (f64), ints (i32),
      layout {param: (offset, shape)}, param_names (declaration order)."""

    kernel_name = "augur_gen_pass"

    def __init__(self, model, hypers, data, ref_compat=True, blocks=None):
        """blocks: the schedule, each an HMC block (a list of continuous
        parameter names) or ("DiscGibbs", [name]) for a discrete parameter."""
        self.model = model
        self._disc_sub = None      # (name, index codes, candidate) inside a DiscGibbs kernel
        self.ref_compat = bool(ref_compat)
        self.syms = {}
        self._int_values = {}      # int arrays by name (index bounds checks)
        self._grid_ranges = {}     # grid variable -> extent, in the current term
        self._reals, self._ints = [], []
        self._nreal = self._nint = 0
        self._bind_hypers(hypers)
        self._bind_params()
        self._bind_data(data)
        self._tmp = 0
        # the log joint, then one pass per HMC block over the declarations
        # whose terms mention the block's variables: the block's full
        # conditional (Core/RwCore.hs:58-67 fullCond, one per kernel at
        # Compile/Compile.hs:182-184).  A factor shared by two variables of
        # one block is counted once (the reference's fullConds' multiplies
        # the variables' full conditionals, which counts it twice).
        decls = [d for d in self.model.decls if d.stochastic]
        kernels = [self._emit_kernel(self.kernel_name, decls)]
        self.grid, self.n_slots = self._last_grid, self._last_slots
        self.block_kernels = []   # (kernel name, grid, kind: "HMC" | "DiscGibbs")
        for b, blk in enumerate(blocks or []):
            kind, vars_ = ("HMC", blk) if isinstance(blk, list) else blk
            vs = set(vars_)
            sub = [d for d in decls if d.name in vs or (_free_vars(d.exp, set()) & vs)]
            if kind == "DiscGibbs":
                if len(vars_) != 1 or self.syms.get(vars_[0]) is None or self.syms[vars_[0]].kind != "dparam":
                    _fail(f"DiscGibbs updates one discrete parameter, got {vars_}")
                name = f"augur_gen_disc_b{b}"
                kernels.append(self._emit_disc_kernel(name, vars_[0], sub))
            else:
                for v in vars_:
                    if self.syms.get(v) is None or self.syms[v].kind != "param":
                        _fail(f"HMC moves continuous parameters; {v} is not one")
                name = f"{self.kernel_name}_b{b}"
                kernels.append(self._emit_kernel(name, sub))
            self.block_kernels.append((name, self._last_grid, kind))
        self.source = (f"constexpr int kRefCompat = {1 if self.ref_compat else 0};\n" + PRELUDE
                       + "".join(kernels))

    # ---- binding ----
    def _pack(self, arr, is_int):
        arr = np.asarray(arr)
        if arr.ndim == 2:
            arr = arr.T  # column-major
        flat = np.ascontiguousarray(arr.ravel(), np.int32 if is_int else np.float64)
        if is_int:
            off = self._nint
            self._ints.append(flat)
            self._nint += flat.size
        else:
            off = self._nreal
            self._reals.append(flat)
            self._nreal += flat.size
        return off

    @staticmethod
    def _as_array(v, ty):
        rank = 0
        t = ty
        while t[0] in ("Vec", "Mat"):
            rank += 1 if t[0] == "Vec" else 2
            t = t[1]
        if rank == 0:
            return None, t
        if isinstance(v, np.ndarray) and v.dtype != object:
            a = v
        else:
            rows = [np.asarray(r) for r in v]
            if rank == 2 and len({r.shape for r in rows}) > 1:
                _fail("ragged arrays are outside the generic lowering")
            a = np.stack(rows) if rank == 2 else np.asarray(v)
        if a.ndim != rank:
            _fail(f"expected an array of rank {rank}, got shape {a.shape}")
        return a, t

    def _bind_hypers(self, hypers):
        for name, ty in self.model.hypers:
            v = hypers[name]
            a, base = self._as_array(v, ty)
            is_int = base == ("Int",)
            if a is None:
                self.syms[name] = _Sym("hyper", (), is_int, value=int(v) if is_int else float(v))
            else:
                self.syms[name] = _Sym("hyper", a.shape, is_int, off=self._pack(a, is_int))
                if is_int:
                    self._int_values[name] = np.asarray(a, np.int64)

    def _grid_sizes(self, d):
        sizes = []
        for var, lo, hi in d.grids:
            if self._const(lo) != 0:
                _fail(f"{d.name}: grids must start at 0")
            n = self._const(hi)
            if not isinstance(n, int) or n < 0:
                _fail(f"{d.name}: grid bound must be a non-negative integer constant")
            sizes.append(n)
        return tuple(sizes)

    def _support(self, d):
        """Values of a discrete parameter: Bernoulli 2, Categorical(pmf) the
        length of its pmf (a Real vector hyper-parameter)."""
        if d.exp[1] == "Bernoulli":
            return 2
        a = d.exp[2][0] if len(d.exp[2]) == 1 else None
        s = self.syms.get(a[1]) if a is not None and a[0] == "var" else None
        if s is None or s.kind != "hyper" or len(s.shape) != 1 or s.is_int:
            _fail(f"{d.name} ~ Categorical(pmf): pmf must be a Real vector hyper-parameter")
        if not 1 <= s.shape[0] <= MAX_SUPPORT:
            _fail(f"{d.name}: the generic DiscGibbs takes at most {MAX_SUPPORT} values")
        return s.shape[0]

    def _bind_params(self):
        off = zoff = 0
        self.layout = {}           # continuous: name -> (offset in q, shape)
        self.zlayout = {}          # discrete: name -> (offset in Z, shape, support)
        self.param_names = []
        for d in self.model.params:
            if not d.stochastic:
                _fail(f"deterministic parameter {d.name}: outside the generic lowering")
            if d.exp[0] != "dist" or d.exp[1] not in PARAM_DISTS | DISC_DISTS:
                _fail(f"parameter {d.name}: the generic lowering samples continuous parameters "
                      f"({', '.join(sorted(PARAM_DISTS))}) with HMC and discrete ones "
                      f"({', '.join(sorted(DISC_DISTS))}) with DiscGibbs")
            if len(d.grids) != len(d.idxs) or len(d.idxs) > 2:
                _fail(f"parameter {d.name}: rank 0-2 over its own grid")
            shape = self._grid_sizes(d)
            size = int(np.prod(shape)) if shape else 1
            if d.exp[1] in DISC_DISTS:
                sym = _Sym("dparam", shape, is_int=True, off=zoff)
                sym.support = self._support(d)
                self.syms[d.name] = sym
                self.zlayout[d.name] = (zoff, shape, sym.support)
                zoff += size
            else:
                self.syms[d.name] = _Sym("param", shape, off=off)
                self.layout[d.name] = (off, shape)
                off += size
            self.param_names.append(d.name)
        # kernels hold q and Z in arrays of at least one element
        self.P = max(off, 1)
        self.NZ = max(zoff, 1)
        if self.P > MAX_PARAMS:
            _fail(f"{self.P} parameter coordinates: the generic lowering holds at most {MAX_PARAMS}")

    def _bind_data(self, data):
        for d in self.model.data:
            if not d.stochastic or d.exp[0] != "dist" or d.exp[1] not in DATA_DISTS:
                _fail(f"data {d.name}: supported distributions are {', '.join(sorted(DATA_DISTS))}")
            if len(d.grids) != len(d.idxs) or len(d.idxs) > 2:
                _fail(f"data {d.name}: rank 0-2 over its own grid")
            shape = self._grid_sizes(d)
            is_int = d.exp[1] in INT_DISTS
            v = data[d.name]
            if shape:
                a = np.asarray(v if not (isinstance(v, np.ndarray) and v.dtype == object)
                               else np.stack([np.asarray(r) for r in v]))
                if a.shape != shape:
                    _fail(f"data {d.name}: expected shape {shape}, got {a.shape}")
                self.syms[d.name] = _Sym("data", shape, is_int, off=self._pack(a, is_int))
                if is_int:
                    self._int_values[d.name] = np.asarray(a, np.int64)
            else:
                self.syms[d.name] = _Sym("data", (), is_int, value=int(v) if is_int else float(v))

    def _const(self, e):
        """Compile-time value of an expression over literals and scalar hypers."""
        k = e[0]
        if k == "lit":
            return e[1]
        if k == "var":
            s = self.syms.get(e[1])
            if s is None or s.kind != "hyper" or s.shape != ():
                _fail(f"{e[1]} is not a scalar hyper-parameter")
            return s.value
        if k == "bin":
            a, b = self._const(e[2]), self._const(e[3])
            if e[1] == "/" and isinstance(a, int) and isinstance(b, int):
                return a // b
            return {"+": a + b, "-": a - b, "*": a * b, "/": a / b}[e[1]]
        _fail(f"expected a constant expression, found {e}")

    @property
    def reals(self):
        return np.concatenate(self._reals) if self._reals else np.zeros(1)

    @property
    def ints(self):
        return np.concatenate(self._ints) if self._ints else np.zeros(1, np.int32)

    # ---- code emission helpers ----
    def _new(self, prefix="v"):
        self._tmp += 1
        return f"{prefix}{self._tmp}"

    def _temp(self, lines, expr, prefix="v"):
        name = self._new(prefix)
        lines.append(f"const double {name} = {expr};")
        return name

    def _slot(self, param, coord):
        """Register slot of a constant coordinate, or None (use LDS atomics)."""
        if param not in self.param_slots:
            off, shape = self.layout[param]
            size = int(np.prod(shape)) if shape else 1
            used = sum(int(np.prod(self.layout[p][1])) if self.layout[p][1] else 1
                       for p in self.param_slots)
            if used + size > MAX_REG_SLOTS:
                self.param_slots[param] = None
            else:
                self.param_slots[param] = used
                for j in range(size):
                    self.slots[off + j] = used + j
        base = self.param_slots[param]
        if base is None:
            return None
        return self.slots[coord]

    def _add(self, lines, param, coord, val):
        if isinstance(coord, int):
            s = self._slot(param, coord)
            if s is not None:
                lines.append(f"acc[{s}] += {val};")
                return
        lines.append(f"atomicAdd(&gs[{coord}], {val});")

    # ---- indices ----
    def _range(self, e):
        """Inclusive bounds of an index expression's values."""
        k = e[0]
        if k == "lit" and isinstance(e[1], int):
            return e[1], e[1]
        if k == "var":
            if e[1] in self._grid_ranges:
                return 0, self._grid_ranges[e[1]] - 1
            s = self.syms.get(e[1])
            if s is not None and s.shape == () and s.kind == "dparam":
                return 0, s.support - 1
            if s is not None and s.shape == () and s.is_int and s.kind != "param":
                return s.value, s.value
        if k == "proj" and e[1][0] == "var" and e[1][1] in self.zlayout:
            self._check_proj(e[1][1], e[2])
            return 0, self.syms[e[1][1]].support - 1
        if k == "proj" and e[1][0] == "var" and e[1][1] in self._int_values:
            self._check_proj(e[1][1], e[2])
            vals = self._int_values[e[1][1]]
            return (int(vals.min()), int(vals.max())) if vals.size else (0, -1)
        if k == "bin" and e[1] in "+-*":
            (al, ah), (bl, bh) = self._range(e[2]), self._range(e[3])
            if e[1] == "+":
                return al + bl, ah + bh
            if e[1] == "-":
                return al - bh, ah - bl
            prods = [al * bl, al * bh, ah * bl, ah * bh]
            return min(prods), max(prods)
        _fail(f"unsupported index expression {e}")

    def _check_proj(self, name, idx_exprs):
        """Every index of name[idx...] stays inside its dimension."""
        s = self.syms[name]
        for dim, e in zip(s.shape, idx_exprs):
            lo, hi = self._range(e)
            if lo < 0 or hi >= dim:
                _fail(f"an index of {name} ranges over [{lo}, {hi}], outside [0, {dim})")

    def _index(self, e, env):
        """Integer index expression: a Python int or a C++ int expression."""
        k = e[0]
        if k == "lit":
            if not isinstance(e[1], int):
                _fail(f"index {e[1]} is not an integer")
            return e[1]
        if k == "var":
            if e[1] in env:
                return env[e[1]]
            s = self.syms.get(e[1])
            if s and s.kind == "dparam" and s.shape == ():
                return self._zread(e[1], s, [])
            if s and s.kind == "hyper" and s.shape == () and s.is_int:
                return s.value
            if s and s.kind == "data" and s.shape == () and s.is_int:
                return s.value
            _fail(f"{e[1]} cannot index")
        if k == "proj":
            name = e[1][1] if e[1][0] == "var" else None
            s = self.syms.get(name)
            if s is None or not s.is_int or s.kind == "param" or len(e[2]) != len(s.shape):
                _fail(f"index {e} must be an element of an int array")
            idx = [self._index(i, env) for i in e[2]]
            if s.kind == "dparam":
                return self._zread(name, s, idx)
            return f"I[{self._flat(s, idx)}]"
        if k == "bin" and e[1] in "+-*":
            a, b = self._index(e[2], env), self._index(e[3], env)
            if isinstance(a, int) and isinstance(b, int):
                return {"+": a + b, "-": a - b, "*": a * b}[e[1]]
            return f"({a} {e[1]} {b})"
        _fail(f"unsupported index expression {e}")

    @staticmethod
    def _flat(s, idx):
        """Flat position of an element of an array symbol (column-major for
        rank-2 data, row-major for parameters)."""
        base = s.off
        if len(idx) == 1:
            i = idx[0]
            return base + i if isinstance(i, int) and isinstance(base, int) else f"{base} + ({i})"
        i, j = idx
        if s.kind in ("param", "dparam"):
            if isinstance(i, int) and isinstance(j, int):
                return base + i * s.shape[1] + j
            return f"{base} + ({i}) * {s.shape[1]} + ({j})"
        if isinstance(i, int) and isinstance(j, int):
            return base + j * s.shape[0] + i
        return f"{base} + ({j}) * {s.shape[0]} + ({i})"

    # ---- forward evaluation ----
    def _elem_code(self, s, pos):
        if s.kind == "param":
            return f"qs[{pos}]"
        if s.kind == "dparam":
            return f"(double)Zc[{pos}]"
        return f"(double)I[{pos}]" if s.is_int else f"R[{pos}]"

    def _zread(self, name, s, idx):
        """Int expression of a discrete parameter's element: its current value
        (Z), or inside a DiscGibbs kernel the candidate value at the point
        being drawn."""
        sub = self._disc_sub
        if sub is not None and sub[0] == name:
            if tuple(str(i) for i in idx) != sub[1]:
                _fail(f"{name} is used at another point than the one being drawn")
            return sub[2]
        return f"Zc[{self._flat(s, idx) if idx else s.off}]"

    def _sym_val(self, name, s, idx, lines):
        """Value of name[idx...] (idx may be shorter than the rank: a view)."""
        rank = len(s.shape)
        if len(idx) > rank:
            _fail(f"{name} indexed {len(idx)} times but has rank {rank}")
        if rank == 0:
            if s.kind == "hyper":
                return _Val(_lit(s.value))
            if s.kind == "data":
                return _Val(_lit(s.value))
            if s.kind == "dparam":
                return _Val(self._temp(lines, f"(double){self._zread(name, s, [])}"))
            return _Val(f"qs[{s.off}]", dep=True, coord=s.off, op=("param", name))
        if len(idx) == rank and s.kind == "dparam":
            return _Val(self._temp(lines, f"(double){self._zread(name, s, idx)}"))
        if len(idx) == rank:
            pos = self._flat(s, idx)
            code = self._temp(lines, self._elem_code(s, pos))
            if s.kind == "param":
                return _Val(code, dep=True, coord=pos, op=("param", name))
            return _Val(code)
        # a view of length shape[-1]: the whole vector, or one row of a matrix
        length = s.shape[-1]
        if rank == 1:
            def pos_of(k):
                return self._flat(s, [k])
        else:
            row = idx[0]

            def pos_of(k):
                return self._flat(s, [row, k])
        dep = s.kind == "param"
        return _Val(length=length, dep=dep, op=("param", name) if dep else None,
                    elem=lambda k: self._elem_code(s, pos_of(k)), vcoord=pos_of if dep else None)

    def _fwd(self, e, env, lines):
        k = e[0]
        if k == "lit":
            return _Val(_lit(e[1]))
        if k == "var":
            if e[1] in env:
                return _Val(f"(double)({env[e[1]]})")
            s = self.syms.get(e[1])
            if s is None:
                _fail(f"unknown name {e[1]}")
            return self._sym_val(e[1], s, [], lines)
        if k == "proj":
            if e[1][0] != "var":
                _fail(f"unsupported projection {e}")
            name = e[1][1]
            s = self.syms.get(name)
            if s is None:
                _fail(f"unknown name {name}")
            self._check_proj(name, e[2])
            return self._sym_val(name, s, [self._index(i, env) for i in e[2]], lines)
        if k == "bin":
            a, b = self._fwd(e[2], env, lines), self._fwd(e[3], env, lines)
            if a.is_vec or b.is_vec:
                _fail("arithmetic on vectors: only dotprod takes vector arguments")
            code = self._temp(lines, f"{a.code} {e[1]} {b.code}")
            return _Val(code, dep=a.dep or b.dep, op=("bin", e[1]), kids=(a, b))
        if k == "prim" and e[1] == "sigmoid":
            a = self._fwd(e[2][0], env, lines)
            if a.is_vec:
                _fail("sigmoid of a vector")
            code = self._temp(lines, f"1.0 / (1.0 + exp(-{a.code}))")
            return _Val(code, dep=a.dep, op=("sigmoid",), kids=(a,))
        if k == "prim" and e[1] == "dotprod":
            a, b = self._fwd(e[2][0], env, lines), self._fwd(e[2][1], env, lines)
            if not (a.is_vec and b.is_vec) or a.length != b.length:
                _fail("dotprod takes two vectors of one length")
            name = self._new()
            lines.append(f"double {name} = 0.0;")
            lines.append("#pragma unroll 16")
            lines.append(f"for (int k_ = 0; k_ < {a.length}; ++k_) {name} += {a.elem('k_')} * {b.elem('k_')};")
            return _Val(name, dep=a.dep or b.dep, op=("dotprod",), kids=(a, b))
        _fail(f"unsupported expression {e}")

    # ---- reverse-mode adjoints ----
    def _bwd(self, v, adj, lines):
        if not v.dep:
            return
        op = v.op[0]
        if op == "param":
            self._add(lines, v.op[1], v.coord, adj)
        elif op == "bin":
            a, b = v.kids
            o = v.op[1]
            if o == "+":
                self._bwd(a, adj, lines)
                self._bwd(b, adj, lines)
            elif o == "-":
                self._bwd(a, adj, lines)
                if b.dep:
                    self._bwd(b, self._temp(lines, f"-{adj}", "d"), lines)
            elif o == "*":
                if a.dep:
                    self._bwd(a, self._temp(lines, f"{adj} * {b.code}", "d"), lines)
                if b.dep:
                    self._bwd(b, self._temp(lines, f"{adj} * {a.code}", "d"), lines)
            else:  # d(a / b) = da / b - a db / b^2
                if a.dep:
                    self._bwd(a, self._temp(lines, f"{adj} / {b.code}", "d"), lines)
                if b.dep:
                    self._bwd(b, self._temp(lines, f"-{adj} * {a.code} / ({b.code} * {b.code})", "d"), lines)
        elif op == "sigmoid":
            self._bwd(v.kids[0], self._temp(lines, f"{adj} * {v.code} * (1.0 - {v.code})", "d"), lines)
        elif op == "dotprod":
            a, b = v.kids
            for x, y in ((a, b), (b, a)):
                if x.dep:
                    self._vec_adjoint(x, y, adj, lines)
        else:
            _fail(f"no adjoint for {op}")

    def _vec_adjoint(self, x, y, adj, lines):
        """acc[q(x_k)] += adj * y_k for every k of a parameter view x."""
        param = x.op[1]
        c0 = x.vcoord(0)
        if isinstance(c0, int) and x.length <= UNROLL_CAP and self._slot(param, c0) is not None:
            s0 = self.slots[c0]
            lines.append("#pragma unroll")
            lines.append(f"for (int k_ = 0; k_ < {x.length}; ++k_) acc[{s0} + k_] += {adj} * {y.elem('k_')};")
        else:
            lines.append(f"for (int k_ = 0; k_ < {x.length}; ++k_) "
                         f"atomicAdd(&gs[{x.vcoord('k_')}], {adj} * {y.elem('k_')});")

    # ---- declarations ----
    def _grid_env(self, d):
        """Grid extents of a declaration and its grid variables as C++ int
        expressions of the flattened index f_."""
        sizes = self._grid_sizes(d)
        self._grid_ranges = {g[0]: n for g, n in zip(d.grids, sizes)}
        env = {}
        if len(sizes) == 1:
            env[d.grids[0][0]] = "f_"
        elif len(sizes) == 2:
            env[d.grids[0][0]] = f"(f_ / {sizes[1]})"
            env[d.grids[1][0]] = f"(f_ % {sizes[1]})"
        return sizes, env

    def _term_lines(self, d, env, ll_var, grads):
        """Code adding the term's log density at the point of env to ll_var,
        and (grads) its reverse-mode contributions."""
        lines = []
        dist, args_e = d.exp[1], d.exp[2]
        if len(args_e) != ARITY[dist]:
            _fail(f"{d.name}: {dist} takes {ARITY[dist]} arguments")
        s = self.syms[d.name]
        pt = self._sym_val(d.name, s, [env[g[0]] for g in d.grids], lines)
        if dist == "Categorical":
            # cbits/augur_dist.c:100-104: log pmf[pt]; the pmf is a hyper-parameter
            a = self._fwd(args_e[0], env, lines)
            if not a.is_vec or a.dep:
                _fail(f"{d.name} ~ Categorical(pmf): pmf must be a Real vector hyper-parameter")
            lines.append(f"{ll_var} += log({a.elem(f'(int){pt.code}')});")
            return lines
        pt_code = f"(int){pt.code}" if dist in INT_DISTS else pt.code
        args = [self._fwd(a, env, lines) for a in args_e]
        for a in args:
            if a.is_vec:
                _fail(f"{d.name}: distribution arguments must be scalars")
        fn = "gen_" + dist.lower()
        call = ", ".join([pt_code] + [a.code for a in args])
        lines.append(f"{ll_var} += {fn}_ll({call});")
        if not grads:
            return lines
        if pt.dep:
            self._bwd(pt, self._temp(lines, f"{fn}_dotpt({call})", "d"), lines)
        for i, a in enumerate(args):
            if a.dep:
                self._bwd(a, self._temp(lines, f"{fn}_dotarg{i + 1}({call})", "d"), lines)
        return lines

    def _term(self, d):
        sizes, env = self._grid_env(d)
        body = "\n      ".join(self._term_lines(d, env, "llacc", True))
        if sizes:
            total = int(np.prod(sizes))
            self.max_range = max(self.max_range, total)
            return (f"  // {d.kind} {d.name} ~ {d.exp[1]}\n"
                    f"  for (int f_ = gtid; f_ < {total}; f_ += gstride) {{\n      {body}\n  }}\n")
        return f"  // {d.kind} {d.name} ~ {d.exp[1]}\n  if (gtid == 0) {{\n      {body}\n  }}\n"

    def _check_aligned(self, zname, d, zsizes):
        """Every mention of zname in d is zname[d's own grid variables], over
        the same extents: the full conditional of one point of zname then
        holds only terms at that point."""
        gv = [g[0] for g in d.grids]

        def walk(e):
            k = e[0]
            if k == "var" and e[1] == zname and zsizes:
                _fail(f"{d.name} uses {zname} whole; DiscGibbs needs it at the term's own point")
            if k == "proj" and e[1] == ("var", zname):
                if [i for i in e[2]] != [("var", v) for v in gv]:
                    _fail(f"{d.name} uses {zname} at another point than its own")
            if k == "proj":
                if e[1] != ("var", zname):
                    walk(e[1])
                for i in e[2]:
                    walk(i)
            elif k in ("prim", "dist"):
                for a in e[2]:
                    walk(a)
            elif k == "bin":
                walk(e[2])
                walk(e[3])
        walk(d.exp)
        if d.name != zname and self._grid_sizes(d) != zsizes:
            _fail(f"{d.name}: DiscGibbs on {zname} needs the terms that use it over its grid")

    def _emit_disc_kernel(self, name, zname, decls):
        """DiscGibbs on a discrete parameter (Core/CgDisc.hs:120-165): at every
        point of its grid, the log full conditional of each value (its prior
        and the terms that use it at that point), then one draw by inverse
        CDF in log space with the maximum subtracted first (the reference
        multiplies pdfs in linear space) from the uniform of Philox counter
        (point, 0, sweep, tag) under key (seed, chain)."""
        zs = self.syms[zname]
        zsizes = zs.shape
        K = zs.support
        for d in decls:
            self._check_aligned(zname, d, zsizes)
        total = int(np.prod(zsizes)) if zsizes else 1
        codes = tuple(self._grid_env(self.model.decl(zname))[1][g[0]]
                      for g in self.model.decl(zname).grids)
        self._disc_sub = (zname, codes, "k_")
        body = []
        try:
            for d in decls:
                _, env = self._grid_env(d)
                body.append(f"// {d.kind} {d.name} ~ {d.exp[1]}")
                body.extend(self._term_lines(d, env, "llk", False))
        finally:
            self._disc_sub = None
        body_code = "\n        ".join(body)
        zpos = self._flat(zs, list(codes)) if codes else zs.off
        g = -(-total // (THREADS * ROWS_PER_THREAD))
        self._last_grid = max(8, min(MAX_GRID, -(-g // 8) * 8))
        self._last_slots = 0
        return (
            f'extern "C" __global__ __launch_bounds__({THREADS}) void {name}(\n'
            f"    const double* __restrict__ q, int* __restrict__ Z, const double* __restrict__ R,\n"
            f"    const int* __restrict__ I, unsigned seed, unsigned sweep, unsigned tag) {{\n"
            f"  constexpr int P = {self.P}, NZ = {self.NZ}, K = {K};\n"
            f"  __shared__ double qs[P];\n"
            + self._mapping_code() +
            f"  for (int j = tid; j < P; j += {THREADS}) qs[j] = q[(size_t)c * P + j];\n"
            f"  __syncthreads();\n"
            f"  int* Zc = Z + (size_t)c * NZ;\n"
            f"  for (int f_ = gtid; f_ < {total}; f_ += gstride) {{\n"
            f"    double lp[K];\n"
            f"    double mx = -1.0e308;\n"
            f"    for (int k_ = 0; k_ < K; ++k_) {{\n"
            f"        double llk = 0.0;\n"
            f"        {body_code}\n"
            f"        lp[k_] = llk;\n"
            f"        mx = llk > mx ? llk : mx;\n"
            f"    }}\n"
            f"    double tot = 0.0;\n"
            f"    for (int k_ = 0; k_ < K; ++k_) {{ lp[k_] = exp(lp[k_] - mx); tot += lp[k_]; }}\n"
            f"    const double u = gen_u01(seed, (unsigned)c, (unsigned)f_, sweep, tag) * tot;\n"
            f"    int pick = K - 1;\n"
            f"    double cum = 0.0;\n"
            f"    for (int k_ = 0; k_ < K; ++k_) {{ cum += lp[k_]; if (u < cum) {{ pick = k_; break; }} }}\n"
            f"    Zc[{zpos}] = pick;\n"
            f"  }}\n"
            f"}}\n")

    def _mapping_code(self):
        return (
            f"  // XCD-aware block mapping: workgroups go round-robin over the 8 XCDs\n"
            f"  // in dispatch order, so the chains of one row tile are placed on one\n"
            f"  // XCD and share its L2 (identity when the grid is not a multiple of 8)\n"
            f"  const int lid_ = blockIdx.y * gridDim.x + blockIdx.x;\n"
            f"  const bool xcd_map_ = (gridDim.x & 7) == 0;\n"
            f"  const int w_ = lid_ >> 3;\n"
            f"  const int c = xcd_map_ ? w_ % gridDim.y : blockIdx.y;\n"
            f"  const int tile = xcd_map_ ? (w_ / gridDim.y) * 8 + (lid_ & 7) : blockIdx.x;\n"
            f"  const int tid = threadIdx.x;\n"
            f"  const int gtid = tile * {THREADS} + tid, gstride = gridDim.x * {THREADS};\n")

    def _emit_kernel(self, name, decls):
        """Source of one pass kernel over the terms of `decls`; sets
        _last_grid and _last_slots."""
        self.slots = {}            # coordinate -> register slot
        self.param_slots = {}      # param -> first slot (params held whole in registers)
        self.max_range = 1
        terms = "".join(self._term(d) for d in decls)
        ns = max(1, len(self.slots))
        coord_of = [0] * ns
        for c, s in self.slots.items():
            coord_of[s] = c
        n_slots = len(self.slots)
        # about ROWS_PER_THREAD points of the longest grid per thread, in a
        # multiple of the 8 XCDs (the block mapping below)
        g = -(-self.max_range // (THREADS * ROWS_PER_THREAD))
        self._last_grid = max(8, min(MAX_GRID, -(-g // 8) * 8))
        self._last_slots = n_slots
        P = self.P
        table = ", ".join(str(c) for c in coord_of)
        return (
            f"__constant__ int kSlotCoord_{name}[{ns}] = {{{table}}};\n"
            f'extern "C" __global__ __launch_bounds__({THREADS}) void {name}(\n'
            f"    const double* __restrict__ q, double* __restrict__ part,\n"
            f"    const double* __restrict__ R, const int* __restrict__ I,\n"
            f"    const int* __restrict__ Z) {{\n"
            f"  constexpr int P = {P}, NS = {n_slots}, NZ = {self.NZ};\n"
            f"  __shared__ double qs[P];\n"
            f"  __shared__ double gs[P + 1];\n"
            f"  __shared__ double red[{THREADS // 64}][NS + 1];\n"
            + self._mapping_code() +
            f"  const int* Zc = Z + (size_t)c * NZ;\n"
            f"  for (int j = tid; j < P; j += {THREADS}) qs[j] = q[(size_t)c * P + j];\n"
            f"  for (int j = tid; j <= P; j += {THREADS}) gs[j] = 0.0;\n"
            f"  __syncthreads();\n"
            f"  double acc[NS + 1];\n"
            f"#pragma unroll\n"
            f"  for (int s = 0; s <= NS; ++s) acc[s] = 0.0;\n"
            f"  double llacc = 0.0;\n"
            f"{terms}"
            f"  acc[NS] = llacc;\n"
            f"#pragma unroll\n"
            f"  for (int s = 0; s <= NS; ++s) {{\n"
            f"    const double w = gen_wave_sum(acc[s]);\n"
            f"    if ((tid & 63) == 0) red[tid >> 6][s] = w;\n"
            f"  }}\n"
            f"  __syncthreads();\n"
            f"  for (int s = tid; s <= NS; s += {THREADS}) {{\n"
            f"    double t = red[0][s];\n"
            f"    for (int w = 1; w < {THREADS // 64}; ++w) t += red[w][s];\n"
            f"    gs[s < NS ? kSlotCoord_{name}[s] : P] += t;\n"
            f"  }}\n"
            f"  __syncthreads();\n"
            f"  double* out = part + ((size_t)c * gridDim.x + tile) * (P + 1);\n"
            f"  for (int j = tid; j <= P; j += {THREADS}) out[j] = gs[j];\n"
            f"}}\n")
