"""Benchmark: min-ESS/sec of the HMC sampler for hierarchical logistic
regression (N = 100k rows, D = 64, 4 chains, HMC [v] (*) HMC [beta] with 20
leapfrog steps of 0.005) on one MI355X, at R-hat < 1.01; beside it LDA
tokens/sec (K = 10, V = 1000, D = 200, ~20k tokens, 100 sweeps, 1 chain).

A step is one sweep of the schedule (one HMC iteration of every block on
every chain).  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_hlr(steps, warmup, seed=2024):
    from augurv2_amd import diagnostics, runtime, workloads
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
    log(f"hlr: warmup {warmup} sweeps")
    s.step(warmup)
    s.trace("state", steps)
    # kernel-time events on every pass launch of one sweep in fifty: about
    # 3,700 timed passes in the default run, few sweeps perturbed.  Sweeps
    # whose global index is a multiple of the period are timed, so a period
    # of at most `steps` always catches one.
    s.profile(min(50, steps))
    log(f"hlr: timed {steps} sweeps")
    t0 = time.perf_counter()
    s.step(steps)  # returns after the stream is synchronised
    t1 = time.perf_counter()
    avg_us, n_timed = s.profile_read()
    draws = s.read_trace("state")  # (steps, C, D + 2)
    keep = list(range(D)) + [D + 1]  # theta and v (no intercept in this model)
    summ = diagnostics.summarize(draws[:, :, keep])
    acc = s.acceptance()
    s.close()
    elapsed = t1 - t0
    min_ess = float(np.min(summ["ess_bulk"]))
    return {
        "elapsed": elapsed, "min_ess": min_ess, "rhat_max": float(np.max(summ["rhat"])),
        "min_ess_tail": float(np.min(summ["ess_tail"])),
        "accept": [float(a) for a in acc.mean(axis=1)], "pass_avg_us": avg_us,
        "pass_launches_timed": n_timed, "N": N, "D": D, "C": C, "L": L, "eps": EPS,
        "x": x, "y": y, "blocks": blocks, "post_mean_v": float(summ["mean"][-1]),
    }


def _min_ess(draws, keep):
    from augurv2_amd import diagnostics
    summ = diagnostics.summarize(draws[:, :, keep])
    return float(np.min(summ["ess_bulk"])), float(np.max(summ["rhat"]))


def run_surrogate(iters=5000, warmup=1000, train=2000, seed=77):
    """HLR N = 1M, D = 128, 8 chains: plain HMC and surrogate-gradient HMC
    (linear least squares + MLP 128-256-256-128 on its residual, trained on
    the pairs of `train` exact iterations after warmup), both with 30
    leapfrog steps of 0.005 (step jittered as the reference does) and the
    exact FP64 Metropolis test; `iters` timed iterations each."""
    from augurv2_amd import runtime, workloads
    from augurv2_amd._lib import HMC_BLOCK_THETA, MWG_BLOCK_V
    N, D, C, L, EPS = 1_000_000, 128, 8, 30, 0.005
    x, y, _, _ = workloads.hlr_data(N, D, seed=0)
    # MWG [v] ~ Normal(v, 1.0) (examples/hlr.py:57): HMC on v stalls for good
    # at small v here (its curvature ~ S / v^3 leaves the step unstable)
    blocks = [(MWG_BLOCK_V, 0.0, 1.0, 0), (HMC_BLOCK_THETA, L * EPS, EPS, -L)]
    keep = list(range(D)) + [D + 1]
    out = {}
    for mode in ("plain", "surrogate"):
        s = runtime.HlrSampler(x, y, blocks, chains=C, seed=seed, ref_compat=False)
        for c, st in enumerate(workloads.hlr_init(D, C, seed=11)):
            s.set("state", st, chain=c)
        log(f"surrogate leg ({mode}): warmup {warmup}")
        s.step(warmup)
        t_fit = 0.0
        if mode == "surrogate":
            s.enable_surrogate(1, hidden=256, train_sweeps=train, max_pairs=train * C * L,
                               epochs=10, batch=4096, lr=2e-3, min_accept=0.6, window=100, seed=5)
            t0 = time.perf_counter()
            s.step(train)  # exact HMC collecting pairs, then training
            t_fit = time.perf_counter() - t0
        s.trace("state", iters)
        log(f"surrogate leg ({mode}): timed {iters}")
        t0 = time.perf_counter()
        s.step(iters)
        el = time.perf_counter() - t0
        ess, rhat = _min_ess(s.read_trace("state"), keep)
        r = {"elapsed_s": el, "ms_per_iter": 1e3 * el / iters, "min_ess": ess, "rhat_max": rhat,
             "min_ess_per_s": ess / el, "accept": float(np.nanmean(s.acceptance()[1]))}
        if mode == "surrogate":
            info = s.surrogate_info()
            r.update({"collect_and_train_s": t_fit, "train_ms": info["train_ms"],
                      "pairs": info["pairs"], "loss": info["loss"], "resid_frac": info["resid_frac"],
                      "phase": info["phase"], "fallbacks": info["fallbacks"],
                      "min_ess_per_s_incl_training": ess / (el + t_fit)})
        s.close()
        out[mode] = r
        log(f"surrogate leg ({mode}): {r}")
    out["ratio"] = out["surrogate"]["min_ess_per_s"] / out["plain"]["min_ess_per_s"]
    out["config"] = (f"HLR N={N} D={D}, {C} chains, {L} leapfrog steps of {EPS} (step jittered), "
                     f"warmup {warmup}, surrogate trained on {train} exact iterations, {iters} timed")
    return out


def run_batched(sweeps=1000, warmup=500, chains=4, seed=31):
    """256 HLR datasets (N_i ~ U[1k, 200k], D_i in {16, 32, 64}) in one
    batched sampler (BF16 X, autotuned rows per block, hipGraph sweep), 20
    leapfrog steps of 0.005 with the step jittered, MWG [v] ~ Normal(v, 1.0).
    Up to 4 chains per dataset the data pass runs on the vector units, from
    5 on the matrix cores.  Beside it the plain sampler's time for the same
    sweeps and chains on each dataset alone, measured on eight datasets and
    interpolated linearly in N_i D_i."""
    from augurv2_amd import runtime, workloads
    from augurv2_amd._lib import HMC_BLOCK_THETA, MWG_BLOCK_V
    L, EPS = 20, 0.005
    xs, ys = workloads.batched_hlr_data()
    s = runtime.BhlrSampler(xs, ys, chains=chains, seed=seed, v_block=("mwg", 1.0),
                            theta_block=(L * EPS, EPS, -L), rows_per_block=0, use_graph=True)
    for i, x in enumerate(xs):
        for c, st in enumerate(workloads.hlr_init(x.shape[1], chains, seed=100 + i)):
            s.set(f"state:{i}", st, chain=c)
    rows, cand = s.tuning()
    log(f"batched: rows/block {rows}, candidates {cand}; adapting warmup {warmup}")
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
    s.close()
    ess_i, rhat_i = [], []
    for i, x in enumerate(xs):
        o, d = int(s.offsets[i]), x.shape[1]
        e, r = _min_ess(draws[:, :, o:o + d + 2], list(range(d)) + [d + 1])
        ess_i.append(e)
        rhat_i.append(r)
    ess_i = np.array(ess_i)
    # plain sampler, one dataset at a time
    work = np.array([x.shape[0] * x.shape[1] for x in xs], np.float64)
    pick = np.argsort(work)[np.linspace(0, len(xs) - 1, 8).astype(int)]
    t_plain = []
    for i in pick:
        p = runtime.HlrSampler(xs[i], ys[i], [(MWG_BLOCK_V, 0.0, 1.0, 0),
                                              (HMC_BLOCK_THETA, L * EPS, EPS, -L)],
                               chains=chains, seed=seed, ref_compat=False)
        p.step(20)
        t1 = time.perf_counter()
        p.step(100)
        t_plain.append((time.perf_counter() - t1) / 100)
        p.close()
    a, b = np.polyfit(work[pick], np.array(t_plain), 1)[::-1]
    plain_total = float(np.sum(a + b * work)) * sweeps
    return {"datasets": len(xs), "chains": chains, "rows_total": int(sum(x.shape[0] for x in xs)),
            "data_pass": ("matrix cores: BF16 V_MFMA, FP32 accumulation" if chains >= 5
                          else "vector units: register form, BF16 rows, FP32 arithmetic"),
            "rows_per_block": rows, "autotune_us": cand, "sweeps": sweeps, "elapsed_s": el,
            "warmup_s": t_warm, "step_size_range": [float(eps_i.min()), float(eps_i.max())],
            "v_proposal_var_range": [float(vvar_i.min()), float(vvar_i.max())],
            "n_rhat_below_1_01": int(np.sum(np.array(rhat_i) < 1.01)),
            "ms_per_sweep": 1e3 * el / sweeps,
            "sum_dataset_min_ess_per_s": float(ess_i.sum() / el),
            "min_dataset_ess": float(ess_i.min()), "median_dataset_ess": float(np.median(ess_i)),
            "rhat_max": float(np.max(rhat_i)), "accept_theta": float(np.nanmean(acc[1])),
            "plain_sequential_s_est": plain_total, "speedup_vs_plain_sequential": plain_total / el}


def run_gmm(sweeps=500, warmup=20, chains=4, seed=5):
    """Multivariate GMM Gibbs, N = 1M, D = 16, K = 32, 4 chains: sweeps/s."""
    from augurv2_amd import runtime, workloads
    N, D, K = 1_000_000, 16, 32
    y, z_true, _, _, _ = workloads.gmm_data(N, D, K, seed=0)
    s = runtime.GmmSampler(y, np.ones(K), np.zeros(D), 10.0 * np.eye(D), D + 2, np.eye(D),
                           chains=chains, seed=seed)
    rng = np.random.default_rng(1)
    for c in range(chains):
        s.set("pis", np.full(K, 1.0 / K), chain=c)
        s.set("mu", y[rng.choice(N, K, replace=False)].astype(np.float64), chain=c)
        s.set("cov", np.tile(np.eye(D), (K, 1, 1)), chain=c)
        s.set("z", rng.integers(0, K, N).astype(np.int32), chain=c)
    s.step(warmup)
    t0 = time.perf_counter()
    s.step(sweeps)
    el = time.perf_counter() - t0
    ll = s.log_density()
    s.close()
    return {"sweeps": sweeps, "elapsed_s": el, "ms_per_sweep": 1e3 * el / sweeps,
            "points_per_s": N * chains * sweeps / el, "final_log_density": [float(v) for v in ll]}


def run_lda(sweeps=100, warmup=10):
    from augurv2_amd import runtime, workloads
    doc_off, w, alpha, beta, _ = workloads.lda_corpus(K=10, V=1000, D=200, mean_len=100, seed=0)
    s = runtime.LdaSampler(10, doc_off, w, alpha, beta, chains=1, seed=7)
    rng = np.random.default_rng(3)
    s.set("z", rng.integers(0, 10, len(w)).astype(np.int32))
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
    return {"tokens": T, "sweeps": sweeps, "tokens_per_s": T * sweeps / (t1 - t0),
            "sweep_us": 1e6 * (t1 - t0) / sweeps, "zstep_avg_us": zavg}


def cpu_baseline(hlr, min_ess_per_iter, budget_s=12.0):
    """The FP64 oracle (a restatement of the reference's libcpu sampler: same
    schedule, same RNG contract, L + 3 data passes per HMC iteration), one
    thread, on the same data; timed over whole iterations of all chains
    until the budget is spent."""
    from oracle import oracle as orc
    x64 = hlr["x"].astype(np.float64)
    m = orc.Hlr(x64, hlr["y"], prior_kind=0, prior_a=1.0, prior_b=1.0, ref_compat=False)
    from augurv2_amd import workloads
    states = [st.copy() for st in workloads.hlr_init(hlr["D"], hlr["C"])]
    kinds = {0: 0, 2: 2}
    it = 0
    t0 = time.perf_counter()
    while True:
        for c in range(hlr["C"]):
            for bi, (kind, sl, eps, fx) in enumerate(hlr["blocks"]):
                m.hmc_step(states[c], kinds[kind], bi, sl, eps, fx, 2024, c, it)
        it += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    el = time.perf_counter() - t0
    return {"value": min_ess_per_iter * it / el, "unit": "min-ESS/s", "cores": 1, "kind": "port",
            "sample": f"{it} HMC sweeps x {hlr['C']} chains of the same workload on the FP64 "
                      f"oracle in {el:.1f} s ({el / it:.2f} s/sweep); min-ESS per sweep taken "
                      "from the GPU run (same algorithm and RNG streams)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 8000 timed sweeps reach R-hat < 1.01 on this workload (2000 reach ~1.04)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--only", default="", help="comma list of legs to run besides the headline "
                    "HLR leg: lda, surrogate, batched, gmm (default: all)")
    ap.add_argument("--batched-sweeps", type=int, default=1000,
                    help="timed sweeps of the batched legs (warmup: half as many)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_mib:
        copy_probe(args.copy_mib)
        return
    legs = set(args.only.split(",")) if args.only else {"lda", "surrogate", "batched", "gmm"}
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the sampler "
              "runtime runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)
    hlr = run_hlr(args.steps, args.warmup)
    log(f"hlr: {hlr['elapsed']:.3f} s, min ESS {hlr['min_ess']:.1f}, R-hat {hlr['rhat_max']:.4f}")
    lda = run_lda() if "lda" in legs else None
    if lda:
        log(f"lda: {lda['tokens_per_s']:.3e} tokens/s")
    surr = run_surrogate() if "surrogate" in legs else None
    bs = max(1, args.batched_sweeps)
    bat = run_batched(sweeps=bs, warmup=max(1, bs // 2)) if "batched" in legs else None
    if bat:
        log(f"batched: {bat}")
    # 16 chains per dataset: the chains-by-rows products on the matrix cores
    bat16 = run_batched(sweeps=bs, warmup=max(1, bs // 2), chains=16) if "batched" in legs else None
    if bat16:
        log(f"batched (16 chains): {bat16}")
    gmm = run_gmm() if "gmm" in legs else None
    if gmm:
        log(f"gmm: {gmm}")
    value = hlr["min_ess"] / hlr["elapsed"]
    N, D = hlr["N"], hlr["D"]
    bytes_per_pass = N * D * 4 + N * 4  # X (FP32) + y (int32): the algorithmic bytes
    if hlr["pass_launches_timed"] == 0:
        raise SystemExit("bench: no pass launch was timed (steps must be at least 1)")
    achieved = bytes_per_pass / (hlr["pass_avg_us"] * 1e-6) / 1e9
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(hlr, hlr["min_ess"] / args.steps)
        log(f"cpu: {cpu['sample']}")
    out = {
        "metric": "min-ESS/sec on 1 GPU (HLR HMC, R-hat<1.01); LDA tokens/sec",
        "value": value,
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
                               "with 20 leapfrog steps of 0.005 (jittered as the reference), FP32 X",
                   "global_batch": N, "chains": 4, "parallelism": "single GPU"},
        "rhat_max": hlr["rhat_max"],
        "min_ess": hlr["min_ess"],
        "min_ess_tail": hlr["min_ess_tail"],
        "accept_rate": hlr["accept"],
        "rhat_ok": hlr["rhat_max"] < 1.01,
        "lda_tokens_per_s": lda["tokens_per_s"] if lda else None,
        "lda": {k: lda[k] for k in ("tokens", "sweeps", "sweep_us", "zstep_avg_us")} if lda else None,
        "surrogate": surr,
        "batched": bat,
        "batched_16_chains": bat16,
        "gmm": gmm,
        "roofline": {"kernel": "k_hlr_pass_r<64,4>", "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": 1024.0 * (PASS_FETCH_KIB * FETCH_SCALE + PASS_WRITE_KIB * WRITE_SCALE),
                     "avg_us": hlr["pass_avg_us"],
                     "launches_timed": hlr["pass_launches_timed"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
