"""Drop-in of pyaugur's AugurOpt / AugurInfer (pyaugur/pyaugur/augurlib.py:
34-222 of the reference) over the MI355X sampler runtime.

Where the reference compiles the model to C/CUDA with its Haskell compiler,
writes, builds and dlopens the generated library, this version parses the
model and schedule in Python, matches them to one of the runtime's sampler
families and creates that sampler through the C-ABI (include/augur_hip.h).
Method names, arguments, the sample dictionary format and the error types
follow the reference:

    with AugurInfer('config.yml', model_text) as infer:
        infer.set_compile_opt(AugurOpt(target='gpu'))
        infer.set_user_sched('ConjGibbs [theta] (*) ConjGibbs [phi] (*) DiscGibbs [z]')
        infer.compile(*hypers)(*data)
        samples = infer.samplen(burnIn=100, numSamples=1000)

Additions (keyword arguments with reference-compatible defaults): chains,
seed, hmc_jitter (False: fixed leapfrog count round(simLen / eps)),
ref_compat (the reference's normal_ll / normal_dotarg2 forms), generic
(True: skip the hand-written families and lower the model generically).
"""
import logging
import os

import numpy as np

from . import augurdistlib as adl
from .augurdistlib import AugurDefault, AugurInitStrat, AugurRandom, AugurUser  # noqa: F401
from .plan import UnsupportedModel, match
from .rvparse import RvParseError, parse_model, parse_schedule

log = logging.getLogger("augurv2_amd")


class AugurOpt:
    """cached: accepted for compatibility (nothing is generated or cached);
    target: 'gpu' runs on the MI355X; 'cpu' (the reference's default, its
    libcpu) is not provided by this runtime and compile raises for it;
    initStrat: initialisation strategy; paramScale: accepted for
    compatibility (unused by the reference's compiler too)."""

    def __init__(self, cached=False, target="cpu", initStrat=None, paramScale=None, chains=1,
                 seed=0, hmc_jitter=True, ref_compat=True, generic=False):
        assert target in ("cpu", "gpu")
        self.cached = cached
        self.target = target
        self.initStrat = initStrat if initStrat is not None else AugurRandom()
        assert isinstance(self.initStrat, AugurInitStrat)
        self.paramScale = paramScale
        self.chains = int(chains)
        self.seed = int(seed)
        self.hmc_jitter = bool(hmc_jitter)
        self.ref_compat = bool(ref_compat)
        self.generic = bool(generic)


def _read_config(path):
    if path is None or not os.path.exists(str(path)):
        return {}
    import yaml
    with open(path) as f:
        cfg = yaml.safe_load(f) or {}
    if "os" in cfg:
        assert cfg["os"] in ("osx", "linux")
    return cfg


class AugurInfer:
    def __init__(self, path_config, model):
        self.model = model
        self.config = _read_config(path_config)  # os, augurdir, ... are not needed here
        self.aopt = None
        self.userSched = None
        self.plan = None
        self.sampler = None
        self.param_names = []
        self.mapSamp = None
        self.mapLL = float("-inf")

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc_value, traceback):
        self.close()

    def close(self):
        if self.sampler is not None:
            self.sampler.close()
            self.sampler = None

    # ---- public interface (augurlib.py:149-222) ----
    def set_compile_opt(self, aopt):
        self.aopt = aopt
        return self

    def set_user_sched(self, userSched):
        self.userSched = userSched
        return self

    def set_pt(self, pt, chain=-1):
        for p in self.param_names:
            assert p in pt
        self.plan.set_point(pt, chain)

    def compile(self, *hypers):
        def _compile(*data):
            if self.aopt is None:
                raise NameError("Set compiler options first...")
            if self.aopt.target != "gpu":
                raise UnsupportedModel("Error: target 'cpu' (the reference's libcpu) is not part "
                                       "of this runtime; use target='gpu'")
            try:
                model = parse_model(self.model)
                sched = parse_schedule(self.userSched) if self.userSched else None
            except RvParseError as e:
                raise NameError("Error: " + str(e)) from e
            if len(hypers) != len(model.hypers):
                raise NameError(f"Error: the model takes {len(model.hypers)} hyper-parameters, "
                                f"got {len(hypers)}")
            if len(data) != len(model.data):
                raise NameError(f"Error: the model has {len(model.data)} data declarations, "
                                f"got {len(data)}")
            hyp = {name: val for (name, _), val in zip(model.hypers, hypers)}
            dat = {d.name: val for d, val in zip(model.data, data)}
            self.plan = match(model, sched, generic=self.aopt.generic)
            self.sampler = self.plan.build(hyp, dat, self.aopt)
            self.param_names = list(self.plan.param_names)
            self._init_chains()
        return _compile

    def _init_chains(self):
        strat = self.aopt.initStrat
        for c in range(self.aopt.chains):
            rng = np.random.default_rng([self.aopt.seed, c])
            if isinstance(strat, AugurUser):
                pt = strat.initPt
            else:
                pt = self.plan.init_point(strat, rng)
            self.plan.set_point(pt, c)

    def curr_ll(self):
        assert self.sampler is not None
        ll = self.sampler.log_density()
        return float(ll[0]) if self.aopt.chains == 1 else ll

    def _point(self):
        if self.aopt.chains == 1:
            return self.plan.get_point(0)
        pts = [self.plan.get_point(c) for c in range(self.aopt.chains)]
        return {k: [p[k] for p in pts] for k in pts[0]}

    def sample(self):
        assert self.sampler is not None
        self.sampler.step(1)
        return self._point()

    def samplen(self, burnIn=0, numSamples=1, thin=1, log_like=False):
        """Returns {name: [value per kept sample]} (a list of per-chain values
        per sample when chains > 1).  As the reference (augurlib.py:195-222):
        numSamples * thin sweeps after burnIn, keeping the state after sweeps
        1, thin + 1, 2 thin + 1, ... (sweep index i kept when i % thin == 0)."""
        assert self.sampler is not None
        assert thin >= 1
        if burnIn:
            self.sampler.step(burnIn)
        samples = {name: [] for name in self.param_names}
        for j in range(numSamples):
            # sweep i = j * thin is the kept one; the thin - 1 after it follow
            self.sampler.step(1)
            if (j * thin) % 100 == 0:
                log.info("Doing iter: %d / %d", j * thin, numSamples)
            pt = self._point()
            if log_like:
                print("log-like[%d]: %s" % (j, self.curr_ll()))
            for name in self.param_names:
                samples[name].append(pt[name])
            if thin > 1:
                self.sampler.step(thin - 1)
        return samples


__all__ = ["AugurOpt", "AugurInfer", "AugurRandom", "AugurDefault", "AugurUser",
           "UnsupportedModel", "adl"]
