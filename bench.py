"""Benchmark: min-ESS/sec of the HMC sampler for hierarchical logistic
regression (N = 100k rows, D = 64, 4 chains, HMC [v] (*) HMC [beta] with 20
leapfrog steps of 0.005) on one MI355X, at R-hat < 1.01; beside it LDA
tokens/sec (K = 10, V = 1000, D = 200, ~20k tokens, 100 sweeps, 1 chain),
the GMM (N = 1M, D = 16, K = 32, 4 chains, 500 sweeps), the batched HLR
sampler (256 datasets) and surrogate-gradient HMC against the plain GPU
sampler (N = 1M, D = 128, 8 chains).

A step of the headline is a block of SWEEPS_PER_STEP sweeps of the schedule
(one HMC iteration of every block on every chain per sweep), so that
`--steps 20 --warmup 5` times 10,000 sweeps after 2,500 warmup sweeps: the
sweeps R-hat < 1.01 needs on this workload.  `value` is null unless every
monitored coordinate has R-hat < 1.01.  Prints one JSON line.
"""
import argparse
import contextlib
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SWEEPS_PER_STEP = 500
RHAT_MAX = 1.01

# HBM bytes per launch of the headline pass kernel: rocprofv3 --pmc FETCH_SIZE
# and WRITE_SIZE (KiB per dispatch) on this workload, each scaled by what the
# counter reads for a plain 256 MiB device copy (`--copy-mib 256`: fetch x2.0,
# write x1.0); profiles/k_hlr_pass_r_hlr_n100k_d64_c4.md
PASS_FETCH_KIB, PASS_WRITE_KIB = 12762.9, 1056.0
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0


def copy_probe(mib):
    """A plain device copy of `mib` MiB, five times: the calibration kernel
    for the traffic counters."""
    import torch
    a = torch.ones(mib * 1024 * 1024 // 4, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copied 5 x {mib} MiB")


def stream_bandwidth(mib=1024, reps=20):
    """Attainable HBM bandwidth on this device: a read and a copy of a
    buffer four times the 256 MB Infinity Cache (GB/s)."""
    import ctypes as C
    from augurv2_amd import _lib
    rd, cp = C.c_double(0.0), C.c_double(0.0)
    _lib.call("augur_op_stream_bandwidth", mib << 20, reps, C.byref(rd), C.byref(cp))
    return {"read_gbs": rd.value, "copy_gbs": cp.value, "buffer_mib": mib}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ---------------------------------------------------------------- CPU side

def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


@contextlib.contextmanager
def one_core():
    """Pins this process to one core for the duration (the `taskset -c`
    protocol of the CPU baseline), then restores the affinity."""
    old = os.sched_getaffinity(0)
    core = min(old)
    os.sched_setaffinity(0, {core})
    try:
        yield core
    finally:
        os.sched_setaffinity(0, old)


def median_time(fn, runs=3):
    """Median wall time of `runs` calls of fn() on one pinned core."""
    times = []
    with one_core() as core:
        for _ in range(runs):
            t0 = time.perf_counter()
            fn()
            times.append(time.perf_counter() - t0)
    return statistics.median(times), times, core


def cpu_record(value, unit, sample, times, core):
    return {"value": value, "unit": unit, "cores": 1, "kind": "port", "sample": sample,
            "cpu_model": cpu_model(), "nproc": os.cpu_count(), "pinned_core": core,
            "runs_s": [round(t, 4) for t in times], "statistic": "median of 3"}


# ---------------------------------------------------------------- legs

def _summary(draws, keep=None):
    from augurv2_amd import diagnostics
    d = draws if keep is None else draws[:, :, keep]
    summ = diagnostics.summarize(d)
    return float(np.min(summ["ess_bulk"])), float(np.max(summ["rhat"])), summ


def _dataset_summaries(draws, offsets, dims):
    """Per dataset of a batched trace (sweeps, chains, sum (D_i + 2)): min
    bulk-ESS and max R-hat over theta and v, and the per-coordinate means and
    MCSEs, computed on a pool of threads (numpy releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor

    def one(i):
        o, d = int(offsets[i]), int(dims[i])
        return _summary(draws[:, :, o:o + d + 2], list(range(d)) + [d + 1])

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        res = list(ex.map(one, range(len(dims))))
    ess = np.array([r[0] for r in res])
    rhat = np.array([r[1] for r in res])
    return ess, rhat, [r[2] for r in res]


def run_hlr(steps, warmup, seed=2024):
    from augurv2_amd import runtime, workloads
    from augurv2_amd._lib import HMC_BLOCK_THETA, HMC_BLOCK_V
    N, D, C, L, EPS = 100_000, 64, 4, 20, 0.005
    x, y, v_true, beta_true = workloads.hlr_data(N, D, seed=0)
    # the reference's HMC (cbits/augur_mcmclib.h:401-405): step jittered in
    # [0.8, 1.2] eps, round(U[0.9, 1.1] * sim_len / step) steps, 20 of 0.005
    # nominal.  Without the jitter the trajectory length 0.1 is a full period
    # (omega T = 2 pi) of the posterior's directions with curvature ~3900 and
    # those directions barely move.
    blocks = [(HMC_BLOCK_V, 1.0, 0.05, 0), (HMC_BLOCK_THETA, L * EPS, EPS, 0)]
    # textbook Normal density: with the reference's normal_dotarg2 (-0.5 * var,
    # cbits/augur_dist.c:393) the HMC force on v is wrong by ~50 at the
    # posterior mode and the v block accepts almost nothing
    s = runtime.HlrSampler(x, y, blocks, chains=C, seed=seed, prior_kind=0, prior_a=1.0,
                           prior_b=1.0, ref_compat=False)
    for c, st in enumerate(workloads.hlr_init(D, C)):
        s.set("state", st, chain=c)
    n_warm, n_timed = warmup * SWEEPS_PER_STEP, steps * SWEEPS_PER_STEP
    log(f"hlr: warmup {warmup} steps ({n_warm} sweeps)")
    s.step(n_warm)
    s.trace("state", n_timed)
    # kernel-time events on every pass launch of one sweep in fifty
    s.profile(50)
    log(f"hlr: timed {steps} steps ({n_timed} sweeps)")
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step(SWEEPS_PER_STEP)  # returns after the stream is synchronised
    t1 = time.perf_counter()
    avg_us, n_prof = s.profile_read()
    final = [s.get("state", c) for c in range(C)]
    draws = s.read_trace("state")  # (sweeps, C, D + 2)
    keep = list(range(D)) + [D + 1]  # theta and v (no intercept in this model)
    min_ess, rhat, summ = _summary(draws, keep)
    acc = s.acceptance()
    s.close()
    return {
        "elapsed": t1 - t0, "min_ess": min_ess, "rhat_max": rhat,
        "min_ess_tail": float(np.min(summ["ess_tail"])), "sweeps": n_timed,
        "accept": [float(a) for a in acc.mean(axis=1)], "pass_avg_us": avg_us,
        "pass_launches_timed": n_prof, "N": N, "D": D, "C": C, "L": L, "eps": EPS,
        "x": x, "y": y, "blocks": blocks, "post_mean_v": float(summ["mean"][-1]), "final": final,
    }


GEN_HLR = """(K : Int, N : Int, a : Real, b : Real, x : Vec (Vec Real)) => {
  param v ~ InvGamma(a, b) ;
  param theta[k] ~ Normal(0.0, v)
      for k <- 0 until K ;
  data y[n] ~ Bernoulli( sigmoid(dotprod(theta, x[n])) )
      for n <- 0 until N ;
}
"""


def run_generic(hlr, sweeps=10_000, warmup=300, seed=77):
    """The headline model and schedule through the generic lowering (the
    data pass generated from the model text, FP64 X, compiled with hipRTC):
    what a model outside the hand-written families gets.  The chains start
    from the headline run's final states, with a seed of their own (chains
    driven by the same Philox streams as the headline's would couple to
    them)."""
    from augurv2_amd import runtime
    from augurv2_amd.pyaugur.lower import Lowering
    from augurv2_amd.pyaugur.rvparse import parse_model
    N, D, C = hlr["N"], hlr["D"], hlr["C"]
    x = hlr["x"].astype(np.float64)
    t0 = time.perf_counter()
    lw = Lowering(parse_model(GEN_HLR), dict(K=D, N=N, a=1.0, b=1.0, x=x), dict(y=hlr["y"]),
                  ref_compat=False, blocks=[["v"], ["theta"]])
    s = runtime.GenSampler(lw, [(["v"], 1.0, 0.05, 0), (["theta"], hlr["L"] * hlr["eps"], hlr["eps"], 0)],
                           chains=C, seed=seed)
    compile_s = time.perf_counter() - t0
    for c, st in enumerate(hlr["final"]):
        # the HLR state is (theta, b, v); the generic vector is (v, theta)
        s.set("state", np.concatenate([[st[D + 1]], st[:D]]), chain=c)
    log(f"generic: warmup {warmup} sweeps")
    s.step(warmup)
    s.trace("state", sweeps)
    s.profile(50)
    t0 = time.perf_counter()
    s.step(sweeps)
    elapsed = time.perf_counter() - t0
    avg_us, n_prof = s.profile_read()
    draws = s.read_trace("state")
    min_ess, rhat, _ = _summary(draws)
    acc = s.acceptance()
    # the data pass alone: the log joint's pass (the theta block's full
    # conditional reads the same data) at the final points
    qf = np.stack([s.get("state", c) for c in range(C)])
    s.profile(1)
    for _ in range(30):
        s.eval(qf)
    data_us, _ = s.profile_read()
    s.close()
    # X (FP64) and y once: the chains of a row tile share them through one
    # XCD's L2
    pass_bytes = N * D * 8 + N * 4
    return {
        "workload": "the headline model and schedule lowered generically (FP64 X), from the "
                    "headline chains' final states",
        "sweeps": sweeps, "warmup": warmup, "elapsed_s": elapsed, "ms_per_sweep": 1e3 * elapsed / sweeps,
        "compile_s": compile_s, "min_ess": min_ess, "rhat_max": rhat, "rhat_ok": rhat < RHAT_MAX,
        "min_ess_per_s": min_ess / elapsed if rhat < RHAT_MAX else None,
        "min_ess_per_s_unchecked": min_ess / elapsed,
        "accept": [float(a) for a in acc.mean(axis=1)], "grid": lw.grid,
        "pass_avg_us_all_blocks": avg_us, "pass_launches_timed": n_prof,
        "data_pass_us": data_us, "data_pass_bytes": pass_bytes,
        "data_pass_gbs": pass_bytes / (data_us * 1e-6) / 1e9 if data_us else None,
    }


def cpu_hlr(hlr, sweeps_per_run=6):
    """The FP64 oracle (a restatement of the reference's libcpu sampler: same
    schedule, same Philox streams, L + 3 data passes per HMC iteration), one
    pinned core, on the same data: median of 3 runs of a truncated number of
    sweeps of all chains, extrapolated linearly per sweep (the protocol for
    long configurations).  The oracle's chain follows the GPU's draw for
    draw (same algorithm and streams), so its min-ESS per sweep is the GPU
    run's."""
    from augurv2_amd import workloads
    from oracle import oracle as orc
    x64 = hlr["x"].astype(np.float64)
    m = orc.Hlr(x64, hlr["y"], prior_kind=0, prior_a=1.0, prior_b=1.0, ref_compat=False)
    kinds = {0: 0, 2: 2}
    start = [st.copy() for st in workloads.hlr_init(hlr["D"], hlr["C"])]

    def run():
        states = [st.copy() for st in start]
        for it in range(sweeps_per_run):
            for c in range(hlr["C"]):
                for bi, (kind, sl, eps, fx) in enumerate(hlr["blocks"]):
                    m.hmc_step(states[c], kinds[kind], bi, sl, eps, fx, 2024, c, it)

    med, times, core = median_time(run)
    s_per_sweep = med / sweeps_per_run
    ok = hlr["rhat_max"] < RHAT_MAX
    ess_per_sweep = hlr["min_ess"] / hlr["sweeps"]
    return cpu_record(ess_per_sweep / s_per_sweep if ok else None, "min-ESS/s",
                      f"{sweeps_per_run} HMC sweeps x {hlr['C']} chains of the same workload per run "
                      f"({s_per_sweep:.3f} s/sweep), extrapolated to the GPU run's min-ESS per sweep "
                      f"({ess_per_sweep:.4f}; same algorithm and Philox streams)"
                      + ("" if ok else "; null: the GPU run did not reach R-hat < 1.01"),
                      times, core)


def run_lda(sweeps=100, warmup=10, cpu=True):
    from augurv2_amd import runtime, workloads
    doc_off, w, alpha, beta, _ = workloads.lda_corpus(K=10, V=1000, D=200, mean_len=100, seed=0)
    s = runtime.LdaSampler(10, doc_off, w, alpha, beta, chains=1, seed=7)
    rng = np.random.default_rng(3)
    z0 = rng.integers(0, 10, len(w)).astype(np.int32)
    s.set("z", z0)
    s.step(warmup)
    t0 = time.perf_counter()
    s.step(sweeps)
    t1 = time.perf_counter()
    # the z step's own duration from kernel-time events, in sweeps after the
    # timed ones (an event pair on every launch would slow the timed sweeps)
    s.profile(1)
    s.step(20)
    zavg, _ = s.profile_read()
    s.close()
    T = int(doc_off[-1])
    # the z step's algorithmic bytes: w read + z write per token
    zbytes = 8.0 * T
    cpu_rec = None
    if cpu:
        # CPU: the oracle's sweep (the reference's libcpu order: theta, phi, z)
        from oracle import oracle as orc
        th, ph = np.zeros((200, 10)), np.zeros((10, 1000))

        def cpu_run():
            z = z0.copy()
            for sw in range(sweeps):
                orc.lda_sweep(doc_off, w, alpha, beta, 7, 0, sw, th, ph, z)

        med, times, core = median_time(cpu_run)
        cpu_rec = cpu_record(T * sweeps / med, "tokens/s",
                             f"the oracle's LDA sweep, {sweeps} sweeps of the same corpus per run",
                             times, core)
    return {"tokens": T, "sweeps": sweeps, "tokens_per_s": T * sweeps / (t1 - t0),
            "sweep_us": 1e6 * (t1 - t0) / sweeps, "zstep_avg_us": zavg,
            "roofline": {"kernel": "k_lda_z", "bound": "hbm", "achieved": zbytes / (zavg * 1e-6) / 1e9,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": zbytes / (zavg * 1e-6) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                         "note": "8 B per token over 19.9k tokens: latency-bound (dependent gathers), "
                                 "not bandwidth-bound, at this size"},
            "cpu_baseline": cpu_rec}


def gmm_summaries(pis, mu, cov):
    """Label-invariant scalars per (sweep, chain): the sorted weights, the
    sorted norms of the means and the sorted traces of the covariances."""
    return np.concatenate([np.sort(pis, axis=-1), np.sort(np.linalg.norm(mu, axis=-1), axis=-1),
                           np.sort(np.trace(cov, axis1=-2, axis2=-1), axis=-1)], axis=-1)


def run_gmm(sweeps=500, warmup=100, chains=4, seed=5, cpu_points=10_000, cpu=True):
    """Multivariate GMM Gibbs, N = 1M, D = 16, K = 32, 4 chains, schedule
    ConjGibbs [pis] (*) ConjGibbs [mu] (*) ConjGibbs [cov] (*) DiscGibbs [z]
    (experiments/experiment.py:127).  Every chain starts from the
    data-generating assignment z (Gibbs from a random start ends in different
    local modes: the data hold components of a few dozen points); the chains
    then run on their own Philox streams.  min-ESS and R-hat are taken over
    label-invariant summaries (sorted weights, mean norms, covariance
    traces)."""
    from augurv2_amd import runtime, workloads
    N, D, K = 1_000_000, 16, 32
    y, z_true, _, _, _ = workloads.gmm_data(N, D, K, seed=0)
    s = runtime.GmmSampler(y, np.ones(K), np.zeros(D), 10.0 * np.eye(D), D + 2, np.eye(D),
                           chains=chains, seed=seed)
    for c in range(chains):
        s.set("pis", np.full(K, 1.0 / K), chain=c)
        s.set("mu", np.zeros((K, D)), chain=c)
        s.set("cov", np.tile(np.eye(D), (K, 1, 1)), chain=c)
        s.set("z", z_true, chain=c)
    s.step(warmup)
    for name in ("pis", "mu", "cov"):
        s.trace(name, sweeps)
    s.profile(25)
    t0 = time.perf_counter()
    s.step(sweeps)
    el = time.perf_counter() - t0
    zavg, _ = s.profile_read()
    summ = gmm_summaries(s.read_trace("pis"), s.read_trace("mu"), s.read_trace("cov"))
    ll = s.log_density()
    s.close()
    min_ess, rhat, _ = _summary(summ)
    ok = rhat < RHAT_MAX
    cpu_rec = None
    if cpu:
        # CPU: the oracle's sweep with the reference's per-(point, component)
        # Cholesky in the categorical step (cbits/augur_dist.c:548-568), on the
        # first cpu_points points, extrapolated linearly in N
        from oracle import oracle as orc
        ys = y[:cpu_points].astype(np.float64)
        g = orc.Gmm(ys, np.ones(K), np.zeros(D), 10.0 * np.eye(D), D + 2, np.eye(D))
        z0 = z_true[:cpu_points].copy()

        def cpu_run():
            pis, mu, cov, z = np.full(K, 1.0 / K), np.zeros((K, D)), np.tile(np.eye(D), (K, 1, 1)), z0.copy()
            g.sweep(pis, mu, cov, z, seed, 0, 0)

        med, times, core = median_time(cpu_run)
        s_sweep_chain = med * N / cpu_points
        ess_per_sweep = min_ess / sweeps
        cpu_rec = cpu_record(ess_per_sweep / (s_sweep_chain * chains) if ok else None, "min-ESS/s",
                             f"one sweep of one chain over the first {cpu_points} points per run "
                             f"({med:.3f} s), extrapolated linearly to N = {N} and {chains} chains "
                             f"({s_sweep_chain * chains:.1f} s per sweep of all chains) at the GPU "
                             "run's min-ESS per sweep", times, core)
    return {"sweeps": sweeps, "warmup": warmup, "chains": chains, "elapsed_s": el,
            "ms_per_sweep": 1e3 * el / sweeps, "points_per_s": N * chains * sweeps / el,
            "min_ess": min_ess, "rhat_max": rhat, "rhat_ok": ok,
            "min_ess_per_s": min_ess / el if ok else None,
            "summary": "sorted pis, sorted |mu_k|, sorted tr(cov_k): 96 label-invariant scalars",
            "zstep_avg_us": zavg, "final_log_density": [float(v) for v in ll],
            "cpu_baseline": cpu_rec}


def run_surrogate(iters=5000, burn=1000, train=1000, seed=77, ablation=True):
    """HLR N = 1M, D = 128, 8 chains, MWG [v] (*) HMC [theta] with 30
    leapfrog steps of 0.005 (step jittered as the reference does), the exact
    FP64 Metropolis test.  Both samplers run burn + train exact iterations
    first (the first 2,000 exact-HMC iterations of the configuration; the
    surrogate records its training pairs during the last `train`), then the
    timed region: `iters` iterations of plain HMC, or the surrogate's training
    followed by `iters` surrogate iterations.  min-ESS over the timed
    iterations."""
    from augurv2_amd import runtime, workloads
    from augurv2_amd._lib import HMC_BLOCK_THETA, MWG_BLOCK_V
    N, D, C, L, EPS = 1_000_000, 128, 8, 30, 0.005
    x, y, _, _ = workloads.hlr_data(N, D, seed=0)
    # MWG [v] ~ Normal(v, 1.0) (examples/hlr.py:57): HMC on v stalls for good
    # at small v here (its curvature ~ S / v^3 leaves the step unstable)
    blocks = [(MWG_BLOCK_V, 0.0, 1.0, 0), (HMC_BLOCK_THETA, L * EPS, EPS, -L)]
    keep = list(range(D)) + [D + 1]
    out = {}
    modes = ("plain", "surrogate", "linear_only") if ablation else ("plain", "surrogate")
    for mode in modes:
        s = runtime.HlrSampler(x, y, blocks, chains=C, seed=seed, ref_compat=False)
        for c, st in enumerate(workloads.hlr_init(D, C, seed=11)):
            s.set("state", st, chain=c)
        log(f"surrogate leg ({mode}): {burn + train} exact iterations")
        t_pre = time.perf_counter()
        s.step(burn)
        if mode != "plain":
            s.enable_surrogate(1, hidden=256, train_sweeps=train, max_pairs=train * C * L,
                               epochs=10, batch=4096, lr=2e-3, min_accept=0.6, window=100, seed=5,
                               linear_only=(mode == "linear_only"))
            s.step(train - 1)  # the last collection sweep (and training) is timed
            extra = 1
        else:
            s.step(train)
            extra = 0
        t_pre = time.perf_counter() - t_pre
        s.trace("state", iters)
        log(f"surrogate leg ({mode}): timed {iters}")
        t0 = time.perf_counter()
        s.step(iters + extra)
        el = time.perf_counter() - t0
        draws = s.read_trace("state")
        ess, rhat, summ = _summary(draws, keep)
        ok = rhat < RHAT_MAX
        r = {"elapsed_s": el, "ms_per_iter": 1e3 * el / iters, "min_ess": ess, "rhat_max": rhat,
             "rhat_ok": ok, "min_ess_per_s": ess / el if ok else None,
             "accept": float(np.nanmean(s.acceptance()[1])), "exact_phase_s": t_pre,
             "whole_job_s": t_pre + el, "post_mean": summ["mean"], "post_mcse": summ["mcse"],
             "post_sd": draws[:, :, keep].reshape(-1, len(keep)).std(axis=0)}
        if mode != "plain":
            info = s.surrogate_info()
            r.update({"train_ms": info["train_ms"], "pairs": info["pairs"], "loss": info["loss"],
                      "resid_frac": info["resid_frac"], "phase": info["phase"],
                      "fallbacks": info["fallbacks"],
                      "min_ess_per_s_excl_training": ess / (el - 1e-3 * info["train_ms"]) if ok else None})
        s.close()
        out[mode] = r
        log(f"surrogate leg ({mode}): " + str({k: v for k, v in r.items()
                                               if not k.startswith("post_")}))
    p, q = out["plain"], out["surrogate"]
    # posterior moments of the surrogate sampler against the FP64 plain
    # sampler: means within 3 combined MCSE (family-wise count beyond 3)
    z = np.abs(q["post_mean"] - p["post_mean"]) / np.hypot(p["post_mcse"], q["post_mcse"])
    res = {"config": (f"HLR N={N} D={D}, {C} chains, {L} leapfrog steps of {EPS} (step jittered), "
                      f"{burn + train} exact iterations first, {iters} timed"),
           "moments_vs_plain": {"max_z": float(z.max()), "n_beyond_3_mcse": int((z > 3).sum()),
                                "n_params": int(z.size),
                                "max_rel_sd_diff": float(np.max(np.abs(q["post_sd"] / p["post_sd"] - 1)))}}
    if p["min_ess_per_s"] and q["min_ess_per_s"]:
        res["ratio_timed_incl_training"] = q["min_ess_per_s"] / p["min_ess_per_s"]
        res["ratio_excl_training"] = q["min_ess_per_s_excl_training"] / p["min_ess_per_s"]
        res["ratio_whole_job"] = (q["min_ess"] / q["whole_job_s"]) / (p["min_ess"] / p["whole_job_s"])
    if "linear_only" in out:
        lo = out["linear_only"]
        res["ablation"] = {"linear_only_accept": lo["accept"], "mlp_accept": q["accept"],
                           "linear_only_min_ess_per_s": lo["min_ess_per_s"],
                           "mlp_min_ess_per_s": q["min_ess_per_s"]}
    for mode, r in out.items():
        res[mode] = {k: v for k, v in r.items() if not k.startswith("post_")}
    if ablation:
        res["ablation_skewed_posterior"] = run_surrogate_skewed()
    return res


def run_surrogate_skewed(sweeps=3000, seed=1):
    """The network's contribution where the linear map falls short: HLR with
    N = 300 rows, D = 8 and coefficients of scale 1.5 (a small, strongly
    informative dataset: a skewed posterior, a gradient far from linear in
    theta).  MWG [v] (*) HMC [theta] with 10 jittered leapfrog steps of 0.06,
    4 chains; exact HMC, the linear surrogate and the network surrogate, each
    300 warmup + 300 collection sweeps, then `sweeps` timed sweeps."""
    from augurv2_amd import _lib, runtime
    from augurv2_amd._lib import HMC_BLOCK_THETA, MWG_BLOCK_V
    rng = np.random.default_rng(seed)
    n, d, C = 300, 8, 4
    x = rng.standard_normal((n, d)).astype(np.float32)
    beta = rng.standard_normal(d) * 1.5
    y = (rng.random(n) < 1 / (1 + np.exp(-(x @ beta)))).astype(np.int32)
    keep = list(range(d)) + [d + 1]
    out = {}
    for mode in ("exact", "linear_only", "mlp"):
        s = runtime.HlrSampler(x, y, [(MWG_BLOCK_V, 0.0, 0.5, 0), (HMC_BLOCK_THETA, 0.6, 0.06, 0)],
                               chains=C, seed=5, ref_compat=False)
        for c in range(C):
            s.set("state", np.concatenate([np.zeros(d), [0.0], [1.0]]), chain=c)
        s.step(300)
        t0 = time.perf_counter()
        if mode != "exact":
            s.enable_surrogate(1, hidden=64, train_sweeps=300, max_pairs=400_000, epochs=60, batch=256,
                               lr=2e-3, min_accept=0.0, window=100_000, seed=3,
                               linear_only=(mode == "linear_only"))
        s.step(300)
        before = np.zeros((2, C, 2))
        _lib.call("augur_stats", s._h, _lib.ptr(before), before.nbytes)
        s.trace("state", sweeps)
        t1 = time.perf_counter()
        s.step(sweeps)
        t2 = time.perf_counter()
        after = np.zeros((2, C, 2))
        _lib.call("augur_stats", s._h, _lib.ptr(after), after.nbytes)
        ess, rhat, _ = _summary(s.read_trace("state"), keep)
        info = s.surrogate_info() if mode != "exact" else {}
        s.close()
        acc = float((after[1, :, 0] - before[1, :, 0]).sum() / (after[1, :, 1] - before[1, :, 1]).sum())
        out[mode] = {"accept": acc, "min_ess": ess, "rhat_max": rhat, "ms_per_sweep": 1e3 * (t2 - t1) / sweeps,
                     "min_ess_per_s": ess / (t2 - t0) if rhat < RHAT_MAX else None,
                     "min_ess_per_s_sampling_only": ess / (t2 - t1) if rhat < RHAT_MAX else None,
                     "network_loss": info.get("loss"), "linear_resid_frac": info.get("resid_frac")}
    out["workload"] = ("HLR N=300, D=8, coefficients ~ N(0, 1.5^2); 10 jittered leapfrog steps of 0.06; "
                       "min-ESS over the timed sweeps, per second from the start of collection (training "
                       "included) and per second of the timed sweeps alone")
    return out


def run_batched(sweeps=2000, warmup=1000, chains=4, seed=31, plain=True, plain_sample=6,
                surrogate_sweeps=0, surrogate_train=200):
    """256 HLR datasets (N_i ~ U[1k, 200k], D_i in {16, 32, 64}) in one
    batched sampler (BF16 X, autotuned rows per block, hipGraph sweep), 20
    leapfrog steps of 0.005 with the step jittered, MWG [v] ~ Normal(v, 1.0);
    step sizes and v proposal variances adapted per dataset in warmup.
    Up to 4 chains per dataset the data pass runs on the vector units, from
    5 on the matrix cores.  Chains start at v = 1, 1.25, 1.5, ... (a draw of
    v from its InvGamma(1, 1) prior can be 10^3, from where a chain random-
    walks down for longer than the warmup).

    With surrogate_sweeps > 0 the same sampler then turns on the linear
    gradient surrogate: surrogate_train exact sweeps collect each dataset's
    (theta, gradient) pairs, the maps are fitted, and surrogate_sweeps
    sweeps run surrogate trajectories with the exact FP64 Metropolis test
    (a dataset whose acceptance falls below 0.6 over 100 sweeps returns to
    exact HMC).  Collection and fit are inside its timed span; its posterior
    means are compared with the exact run's, coordinate by coordinate.

    Beside it the plain sampler (FP32 X, one dataset at a time) on a
    stratified sample of datasets, with each dataset's adapted step size and
    proposal variance: its min-ESS/s per dataset is measured, and the time
    it would take to reach the batched run's min-ESS on every dataset is
    extrapolated from the sample in proportion to N_i D_i."""
    from augurv2_amd import runtime, workloads
    from augurv2_amd._lib import HMC_BLOCK_THETA, MWG_BLOCK_V
    L, EPS = 20, 0.005
    xs, ys = workloads.batched_hlr_data()
    s = runtime.BhlrSampler(xs, ys, chains=chains, seed=seed, v_block=("mwg", 1.0),
                            theta_block=(L * EPS, EPS, -L), rows_per_block=0, use_graph=True)
    for i, x in enumerate(xs):
        for c, st in enumerate(workloads.hlr_init(x.shape[1], chains, seed=100 + i)):
            st[-1] = 1.0 + 0.25 * c
            s.set(f"state:{i}", st, chain=c)
    rows, cand = s.tuning()
    log(f"batched ({chains} chains): rows/block {rows}, candidates {cand}; adapting warmup {warmup}")
    # per-dataset step sizes (dual averaging) and v proposal variances adapt
    # during warmup, then stay fixed: one step size cannot suit datasets of
    # 1k and 200k rows with variances from 0.1 to 100
    t0 = time.perf_counter()
    s.adapt(warmup, 0.8)
    t_warm = time.perf_counter() - t0
    eps_i, vvar_i = s.step_sizes()
    s.trace("state", sweeps)
    t0 = time.perf_counter()
    s.step(sweeps)
    el = time.perf_counter() - t0
    draws = s.read_trace("state")  # (sweeps, C, sum(D_i + 2))
    acc = s.acceptance()
    dims = np.array([x.shape[1] for x in xs])
    srun = None
    if surrogate_sweeps:
        log(f"batched ({chains} chains): surrogate, {surrogate_train} collection sweeps, "
            f"{surrogate_sweeps} timed")
        acc_before = s.acceptance_counts()
        t0 = time.perf_counter()
        s.surrogate(train_sweeps=surrogate_train, ridge=1e-6, min_accept=0.6, window=100)
        s.step(surrogate_train)  # exact sweeps that record the pairs, then the fit
        t_fit = time.perf_counter() - t0
        s.trace("state", surrogate_sweeps)
        s.step(surrogate_sweeps)
        el_s = time.perf_counter() - t0
        srun = {"draws": s.read_trace("state"), "elapsed_s": el_s, "collect_fit_s": t_fit,
                "info": s.surrogate_info(), "modes": s.surrogate_modes(),
                "acc": s.acceptance_counts() - acc_before}
    s.close()
    ess_i, rhat_i, summ_i = _dataset_summaries(draws, s.offsets, dims)
    conv = rhat_i < RHAT_MAX
    out = {"datasets": len(xs), "chains": chains, "rows_total": int(sum(x.shape[0] for x in xs)),
           "data_pass": ("matrix cores: BF16 V_MFMA, FP32 accumulation" if chains >= 5
                         else "vector units: register form, BF16 rows, FP32 arithmetic"),
           "rows_per_block": rows, "autotune_us": cand, "sweeps": sweeps, "warmup": warmup,
           "elapsed_s": el, "warmup_s": t_warm,
           "step_size_range": [float(eps_i.min()), float(eps_i.max())],
           "v_proposal_var_range": [float(vvar_i.min()), float(vvar_i.max())],
           "n_rhat_below_1_01": int(conv.sum()),
           "not_converged": [{"dataset": int(i), "rows": int(xs[i].shape[0]), "d": int(xs[i].shape[1]),
                              "rhat": float(rhat_i[i])} for i in np.nonzero(~conv)[0]],
           "ms_per_sweep": 1e3 * el / sweeps,
           "sum_converged_dataset_min_ess_per_s": float(ess_i[conv].sum() / el),
           "min_converged_dataset_ess": float(ess_i[conv].min()) if conv.any() else None,
           "median_dataset_ess": float(np.median(ess_i)),
           "rhat_max": float(np.max(rhat_i)), "accept_theta": float(np.nanmean(acc[1]))}
    sout = None
    if srun:
        ess_s, rhat_s, summ_s = _dataset_summaries(srun["draws"], s.offsets, dims)
        conv_s = rhat_s < RHAT_MAX
        # acceptance over the collection and surrogate sweeps only
        with np.errstate(invalid="ignore"):
            acc_s = srun["acc"][..., 0] / srun["acc"][..., 1]
        # posterior means against the exact run, datasets converged in both
        zs = []
        for i in np.nonzero(conv & conv_s)[0]:
            zs.append(np.abs(summ_i[i]["mean"] - summ_s[i]["mean"])
                      / np.hypot(summ_i[i]["mcse"], summ_s[i]["mcse"]))
        zs = np.concatenate(zs) if zs else np.zeros(0)
        modes = srun["modes"]
        sout = {"collection_sweeps": surrogate_train, "sweeps": surrogate_sweeps,
                "elapsed_s": srun["elapsed_s"], "collect_fit_s": srun["collect_fit_s"],
                "ms_per_sweep": 1e3 * (srun["elapsed_s"] - srun["collect_fit_s"]) / surrogate_sweeps,
                "fit_ms": srun["info"]["fit_ms"],
                "datasets_on_surrogate": int(modes.sum()),
                "fallback_datasets": [{"dataset": int(i), "rows": int(xs[i].shape[0]),
                                       "d": int(xs[i].shape[1])} for i in np.nonzero(modes == 0)[0]],
                "accept_theta": float(np.nanmean(acc_s[1])),
                "n_rhat_below_1_01": int(conv_s.sum()),
                "not_converged": [{"dataset": int(i), "rows": int(xs[i].shape[0]), "d": int(xs[i].shape[1]),
                                   "rhat": float(rhat_s[i])} for i in np.nonzero(~conv_s)[0]],
                "rhat_max": float(np.max(rhat_s)),
                "sum_converged_dataset_min_ess_per_s": float(ess_s[conv_s].sum() / srun["elapsed_s"]),
                "median_dataset_ess": float(np.median(ess_s)),
                "moments_vs_exact": {
                    "coordinates": int(zs.size),
                    "frac_within_3_mcse": float(np.mean(zs < 3.0)) if zs.size else None,
                    "max_abs_z": float(zs.max()) if zs.size else None,
                    "expected_frac_beyond_3_if_equal": 0.0027,
                    "what": "per coordinate of theta, b, v of every dataset converged in both runs: "
                            "|mean_exact - mean_surrogate| / hypot(MCSE_exact, MCSE_surrogate)"}}
        out["surrogate"] = sout
    if not plain:
        return out
    # the plain sampler, one dataset at a time, on datasets spread over the
    # range of N_i D_i (converged ones), with the batched run's adapted settings
    work = np.array([x.shape[0] * x.shape[1] for x in xs], np.float64)
    order = [i for i in np.argsort(work) if conv[i]]
    pick = [order[j] for j in np.linspace(0, len(order) - 1, plain_sample).astype(int)]
    p_sweeps = max(200, sweeps // 2)
    rows_p = []
    for i in pick:
        p = runtime.HlrSampler(xs[i], ys[i], [(MWG_BLOCK_V, 0.0, float(vvar_i[i]), 0),
                                              (HMC_BLOCK_THETA, L * float(eps_i[i]), float(eps_i[i]), -L)],
                               chains=chains, seed=seed + 1, ref_compat=False)
        for c in range(chains):
            p.set("state", draws[-1, c, int(s.offsets[i]):int(s.offsets[i]) + xs[i].shape[1] + 2], chain=c)
        p.step(100)
        p.trace("state", p_sweeps)
        t1 = time.perf_counter()
        p.step(p_sweeps)
        t_p = time.perf_counter() - t1
        d = xs[i].shape[1]
        e_p, r_p, _ = _summary(p.read_trace("state"), list(range(d)) + [d + 1])
        p.close()
        rows_p.append({"dataset": int(i), "rows": int(xs[i].shape[0]), "d": int(d),
                       "plain_min_ess_per_s": e_p / t_p, "plain_rhat": r_p,
                       "batched_min_ess": float(ess_i[i]),
                       "plain_s_to_batched_ess": float(ess_i[i]) / (e_p / t_p)})
    samp_t = sum(r["plain_s_to_batched_ess"] for r in rows_p)
    seq_est = samp_t * float(work[conv].sum() / work[pick].sum())
    if sout:
        # the same extrapolation to the surrogate run's min-ESS per dataset
        samp_s = sum(float(ess_s[r["dataset"]]) / r["plain_min_ess_per_s"] for r in rows_p)
        seq_s = samp_s * float(work[conv_s].sum() / work[pick].sum())
        sout["plain_sequential_s_to_same_ess_est"] = seq_s
        sout["min_ess_per_s_ratio_vs_plain_sequential"] = seq_s / srun["elapsed_s"]
    out.update({"plain_sample": rows_p, "plain_sweeps": p_sweeps,
                "plain_sequential_s_to_same_ess_est": seq_est,
                "min_ess_per_s_ratio_vs_plain_sequential": seq_est / el,
                "plain_comparison": (f"plain sampler on {len(pick)} converged datasets spread over "
                                     "N_i D_i; time to reach the batched run's min-ESS on each, "
                                     "extrapolated to all converged datasets in proportion to "
                                     "N_i D_i")})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help=f"timed steps of the headline, {SWEEPS_PER_STEP} sweeps each")
    ap.add_argument("--warmup", type=int, default=5, help="untimed warmup steps of the headline")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--only", default="", help="comma list of legs to run besides the headline "
                    "HLR leg: lda, surrogate, batched, gmm, generic (default: all); with "
                    "--skip-headline also batched4 (the 4-chain batched leg alone; batched is then "
                    "the 16-chain one) and stream")
    ap.add_argument("--batched-sweeps", type=int, default=2000,
                    help="timed sweeps of the batched legs (warmup: half as many)")
    ap.add_argument("--batched-surrogate-sweeps", type=int, default=None,
                    help="timed surrogate sweeps after the 4-chain batched leg, half as many after "
                         "the 16-chain one (0: none; default 10000, or 0 with --skip-headline)")
    ap.add_argument("--surrogate-iters", type=int, default=5000,
                    help="timed iterations of each surrogate-leg sampler (shorter for profiling)")
    ap.add_argument("--surrogate-exact", type=int, default=2000,
                    help="exact iterations before the timed ones (half burn-in, half collection)")
    ap.add_argument("--gmm-sweeps", type=int, default=500, help="timed sweeps of the GMM leg")
    ap.add_argument("--generic-sweeps", type=int, default=10_000,
                    help="timed sweeps of the generic-lowering leg")
    ap.add_argument("--skip-headline", action="store_true",
                    help="run only the legs named by --only (profiling runs; no bench line)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_mib:
        copy_probe(args.copy_mib)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the sampler "
              "runtime runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1:
        raise SystemExit("bench: --steps must be at least 1")
    legs = set(args.only.split(",")) if args.only else {"lda", "surrogate", "batched", "gmm", "generic"}
    if args.skip_headline:
        bss_only = max(0, args.batched_surrogate_sweeps or 0)
        for leg, fn in (("lda", lambda: run_lda(cpu=False)),
                        ("gmm", lambda: run_gmm(sweeps=args.gmm_sweeps, cpu=False)),
                        ("surrogate", lambda: run_surrogate(iters=args.surrogate_iters,
                                                            burn=max(1, args.surrogate_exact // 2),
                                                            train=max(1, args.surrogate_exact // 2),
                                                            ablation=False)),
                        ("batched", lambda: run_batched(sweeps=args.batched_sweeps,
                                                        warmup=max(1, args.batched_sweeps // 2),
                                                        chains=16, plain=False,
                                                        surrogate_sweeps=bss_only // 2)),
                        ("batched4", lambda: run_batched(sweeps=args.batched_sweeps,
                                                         warmup=max(1, args.batched_sweeps // 2),
                                                         chains=4, plain=False,
                                                         surrogate_sweeps=bss_only)),
                        ("generic", lambda: run_generic(run_hlr(1, 1), sweeps=args.generic_sweeps,
                                                        warmup=50)),
                        ("stream", lambda: log(str(stream_bandwidth(256, 3))))):
            if leg in legs:
                log(f"{leg}: done ({type(fn()).__name__})")
        return
    hlr = run_hlr(args.steps, args.warmup)
    log(f"hlr: {hlr['elapsed']:.3f} s, min ESS {hlr['min_ess']:.1f}, R-hat {hlr['rhat_max']:.4f}")
    gen = run_generic(hlr, sweeps=args.generic_sweeps) if "generic" in legs else None
    if gen:
        log(f"generic: {gen}")
    lda = run_lda(cpu=not args.no_cpu) if "lda" in legs else None
    if lda:
        log(f"lda: {lda['tokens_per_s']:.3e} tokens/s")
    gmm = run_gmm(sweeps=args.gmm_sweeps, cpu=not args.no_cpu) if "gmm" in legs else None
    if gmm:
        log(f"gmm: {gmm}")
    half = max(1, args.surrogate_exact // 2)
    surr = (run_surrogate(iters=args.surrogate_iters, burn=half, train=half)
            if "surrogate" in legs else None)
    bs = max(1, args.batched_sweeps)
    bss = max(0, 10_000 if args.batched_surrogate_sweeps is None else args.batched_surrogate_sweeps)
    bat = (run_batched(sweeps=bs, warmup=max(1, bs // 2), surrogate_sweeps=bss)
           if "batched" in legs else None)
    if bat:
        log(f"batched: {bat}")
    # 16 chains per dataset: the chains-by-rows products on the matrix cores
    bat16 = (run_batched(sweeps=bs, warmup=max(1, bs // 2), chains=16, surrogate_sweeps=bss // 2)
             if "batched" in legs else None)
    if bat16:
        log(f"batched (16 chains): {bat16}")
    ok = hlr["rhat_max"] < RHAT_MAX
    raw = hlr["min_ess"] / hlr["elapsed"]
    N, D = hlr["N"], hlr["D"]
    bytes_per_pass = N * D * 4 + N * 4  # X (FP32) + y (int32): the algorithmic bytes
    if hlr["pass_launches_timed"] == 0:
        raise SystemExit("bench: no pass launch was timed")
    achieved = bytes_per_pass / (hlr["pass_avg_us"] * 1e-6) / 1e9
    stream = stream_bandwidth()
    log(f"stream: {stream}")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_hlr(hlr)
        log(f"cpu: {cpu['sample']}")
    vs_plain = None
    if surr:
        vs_plain = {k: surr.get(k) for k in ("ratio_timed_incl_training", "ratio_excl_training",
                                             "ratio_whole_job", "config")}
        vs_plain["accelerated_min_ess_per_s"] = surr["surrogate"]["min_ess_per_s"]
        vs_plain["plain_min_ess_per_s"] = surr["plain"]["min_ess_per_s"]
    for name, b in (("batched_plus_surrogate_4_chains", bat), ("batched_plus_surrogate_16_chains", bat16)):
        bsr = b.get("surrogate") if b else None
        if bsr and "min_ess_per_s_ratio_vs_plain_sequential" in bsr:
            vs_plain = vs_plain or {}
            vs_plain[name] = {
                "ratio_vs_plain_sequential": bsr["min_ess_per_s_ratio_vs_plain_sequential"],
                "ratio_exact_batched_vs_plain_sequential": b["min_ess_per_s_ratio_vs_plain_sequential"],
                "datasets_at_rhat_below_1_01": bsr["n_rhat_below_1_01"],
                "moments_frac_within_3_mcse": bsr["moments_vs_exact"]["frac_within_3_mcse"],
                "config": f"256 HLR datasets, {b['chains']} chains each; timed span = "
                          f"{bsr['collection_sweeps']} exact collection sweeps + fit + "
                          f"{bsr['sweeps']} surrogate sweeps; plain = one dataset at a time"}
    out = {
        "metric": "min-ESS/sec on 1 GPU (HLR HMC, R-hat<1.01) vs plain GPU sampler; LDA tokens/sec",
        "value": raw if ok else None,
        "unit": "min-ESS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * hlr["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "HLR HMC: N=100000 rows, D=64, 4 chains, HMC [v] (*) HMC [beta] "
                               "with 20 leapfrog steps of 0.005 (jittered as the reference), FP32 X; "
                               f"one step = {SWEEPS_PER_STEP} sweeps",
                   "global_batch": N, "chains": 4, "sweeps_per_step": SWEEPS_PER_STEP,
                   "parallelism": "single GPU"},
        "rhat_max": hlr["rhat_max"],
        "rhat_ok": ok,
        "min_ess": hlr["min_ess"],
        "min_ess_tail": hlr["min_ess_tail"],
        "min_ess_per_s_unchecked": raw,
        "timed_sweeps": hlr["sweeps"],
        "accept_rate": hlr["accept"],
        "vs_plain_gpu_sampler": vs_plain,
        "lda_tokens_per_s": lda["tokens_per_s"] if lda else None,
        "lda": lda,
        "gmm": gmm,
        "surrogate": surr,
        "batched": bat,
        "batched_16_chains": bat16,
        "generic_lowering": gen,
        "roofline": {"kernel": "k_hlr_pass_r<64,4,false,LL> (LL = false: gradient-only leapfrog "
                               "passes; true: cache and last passes; avg_us over both)",
                     "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": 1024.0 * (PASS_FETCH_KIB * FETCH_SCALE + PASS_WRITE_KIB * WRITE_SCALE),
                     "avg_us": hlr["pass_avg_us"], "launches_timed": hlr["pass_launches_timed"],
                     "note": "X (25.6 MB) stays in the 256 MB Infinity Cache between passes: the "
                             "pass is latency-bound (launch + one memory round), not HBM-bound"},
        "cpu_baseline": cpu,
        "attainable_bandwidth": stream,
    }
    print(json.dumps(out, default=float), flush=True)


if __name__ == "__main__":
    main()
