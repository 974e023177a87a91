"""Benchmark: 1024-chain NUTS on Bayesian logistic regression (N=2M, D=256).

A "step" is one NUTS iteration of every chain (each chain's full trajectory:
momentum draw, iterative tree, multinomial proposal) on the synthetic
2,000,000 x 256 design resident in HBM.  Set-up (untimed): a Laplace
approximation of the posterior (Newton iterations on U in float64) gives
each chain's initial point (the mode plus a uniform(-2, 2) jitter in units
of the Laplace standard deviation, passed as init_params) and the initial
diagonal inverse_mass_matrix (the Laplace variances); the sampler runs with
find_heuristic_step_size=True.  --warmup W untimed iterations follow (dual
averaging + diagonal Welford on the window schedule of W; --mass laplace
keeps the Laplace matrix fixed instead), then K timed sampling iterations.  value = gradient evaluations per second over the timed
region (useful leapfrog steps summed over chains); min-ESS/s is computed
from the K timed draws of every chain.  The BF16-MFMA form is the measured
line; the plain FP32 form (same adapted state, same kernels otherwise) is
timed beside it, and the CPU baseline (the batched potential + gradient of
oracle/libcpu_nuts.so, the C/OpenMP restatement of the same path, on the
host's cores) on a bounded sample.  A run whose timed
transitions mostly diverge (or accept almost nothing) prints its line with
"valid": false and exits with status 3.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAKS = {"bf16": 2500.0, "fp32": 157.3, "fp8": 5000.0}  # dense TFLOP/s, MI355X


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


# Untimed adaptation iterations of the logistic workloads at least: fewer
# leave some of the 1024 chains with unadapted step sizes (R-hat well above
# 1.01 and a long tail of deep trees), whatever --warmup asks for.
MIN_ADAPT = 150

# (timed, warmup) iterations per workload when not given
DEFAULT_ITERS = {"c5": (120, 150), "c2": (1000, 1000), "c1": (500, 500), "c3": (600, 600),
                 "c4": (500, 500), "ecs": (100, 100)}


def run_form(model, C, num_warmup, steps, z0, keys, state=None, collect=True, imm=None,
             heuristic=False, dense=False, adapt_mass=True, regularize=True, launch_shape=None):
    import torch
    from numpyro_amd.infer.hmc import Engine
    eng = Engine(model, C, "NUTS", num_warmup, find_heuristic_step_size=heuristic,
                 dense_mass=dense, adapt_mass_matrix=adapt_mass,
                 regularize_mass_matrix=regularize)
    if launch_shape:
        eng.set_launch_shape(launch_shape)
    eng.init(z0, keys, imm)
    t_start = time.perf_counter()
    eng.set_progress(lambda ticks, pending: log(
        f"  {getattr(model, 'x_dtype', type(model).__name__)}: {ticks} ticks/launches, "
        f"{pending} chains running, "
        f"{time.perf_counter() - t_start:.0f}s"), every_ticks=progress_ticks(model, C))
    warm_s = 0.0
    if state is not None:
        eng.set_state(state)
    elif num_warmup > 0:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.run(num_warmup)
        torch.cuda.synchronize()
        warm_s = time.perf_counter() - t0
    i0 = int(eng.get_state().i.max())
    eng.timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = eng.run(i0 + steps, collect_fields=("z", "num_steps", "accept_prob", "diverging")
                  if collect else ("num_steps",), start_idx=i0, thinning=1, size=steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms, kn = eng.timing(False)
    return eng, out, wall, warm_s, kms, kn


def progress_ticks(model, C):
    # about one report every 20 s at a few ms per tick for the largest designs
    work = model.algorithmic_flops(C)
    return int(max(256, min(65536, 2e16 / max(work, 1.0))))


def min_ess(z):
    """min ESS and max split R-hat over coordinates of draws z [C, S, D],
    computed on the GPU (nax_diagnostics; the reference's formulas)."""
    from numpyro_amd.diagnostics import device_diagnostics, effective_sample_size
    if z.shape[1] > 4096:
        return float(np.min(effective_sample_size(z.cpu().numpy().astype(np.float64)))), \
            float("nan")
    ess, rhat = device_diagnostics(z)
    rhat = rhat.cpu().numpy() if z.shape[1] >= 4 else np.full(z.shape[2], np.nan)
    return float(ess.min()), float(np.nanmax(rhat))


def degenerate(div, acc, transitions):
    """True when more than half of the timed transitions diverged or the
    mean acceptance probability is below 0.05."""
    return div > 0.5 * transitions or not acc >= 0.05


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Host threads the baseline uses: the process's CPU affinity, capped at
    the 16 cores a one-GPU share of the box holds."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(X, y, z0, budget_s, threads, chains=64):
    """The reference's vectorized CPU path does, per leapfrog step of the
    chain batch, one X Z^T and one X^T R over all chains (mcmc.py:720-724,
    vmap).  Times that batched potential + gradient (oracle/libcpu_nuts.so,
    C/OpenMP: every thread one contiguous row block, one pass over each row
    for all chains) for `chains` chains at the chains' initial points,
    repeated for about budget_s seconds.  Returns (grad evals/s, chains,
    calls, seconds)."""
    from oracle import cpu_baseline as cb
    Xh = X.cpu().numpy()
    yh = y.cpu().numpy()
    zc = z0[:chains].cpu().numpy()
    t0 = time.perf_counter()
    cb.logistic_batch_grad(Xh, yh, zc, threads=threads)
    per = time.perf_counter() - t0
    calls = int(max(1, min(200, budget_s / max(per, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(calls):
        cb.logistic_batch_grad(Xh, yh, zc, threads=threads)
    wall = time.perf_counter() - t0
    return chains * calls / wall, chains, calls, wall


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed iterations (default per config: c5 120, c2 1000, c1 500, "
                         "c3 600, c4 500, ecs 100)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed warmup iterations (default per config: c5 150, c2 1000, "
                         "c1 500, c3 600, c4 500, ecs 100)")
    ap.add_argument("--config", default="c5", choices=["c5", "c2", "c1", "c3", "c4", "ecs"])
    ap.add_argument("--subsample", type=int, default=1000, help="ecs: subsample size m")
    ap.add_argument("--num-blocks", type=int, default=100, help="ecs: Gibbs blocks")
    ap.add_argument("--noncentred", action="store_true", help="c3: non-centred SV form")
    ap.add_argument("--per-model", default=None, help="c4: write per-model diagnostics (JSON)")
    ap.add_argument("--launch-shape", type=int, default=None,
                    help="persistent models (c1, c3, c4): fix the launch shape instead of "
                         "autotuning it (waves per chain; c4: chains per workgroup + 32 x "
                         "(waves per chain - 1))")
    ap.add_argument("--x-dtype", default=None, choices=["bf16", "fp32", "fp8"],
                    help="storage form of X (default: bf16 for c5, fp32 for c1 and c2)")
    ap.add_argument("--bf16-kernel", default="single", choices=["single", "two-copy"],
                    help="bf16 storage: one X copy (bf16 residuals, the library default) or "
                         "two (fp16 residuals)")
    ap.add_argument("--rows", type=int, default=None, help="override N (smoke runs)")
    ap.add_argument("--chains", type=int, default=None)
    ap.add_argument("--fp32-steps", type=int, default=24,
                    help="timed iterations of the plain FP32 form (0: skip)")
    ap.add_argument("--cpu-budget", type=float, default=10.0,
                    help="seconds of CPU-baseline work (0: skip)")
    ap.add_argument("--dense-mass", action="store_true",
                    help="logistic workloads: dense_mass=True, initial inverse_mass_matrix "
                         "the Laplace covariance")
    ap.add_argument("--mass", default="adapt", choices=["adapt", "adapt-unregularized", "laplace"],
                    help="logistic workloads: adapt the inverse mass matrix by Welford "
                         "windows starting from the Laplace variances (default), the same "
                         "with regularize_mass_matrix=False, or keep the Laplace matrix "
                         "fixed (adapt_mass_matrix=False)")
    args = ap.parse_args()
    steps, warm = DEFAULT_ITERS[args.config]
    args.steps = steps if args.steps is None else args.steps
    args.warmup = warm if args.warmup is None else args.warmup
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); chains "
              "would shard across GPUs with no communication", file=sys.stderr)
        sys.exit(2)

    import torch
    from numpyro_amd import random as nrandom
    from numpyro_amd import workloads
    from numpyro_amd.models import LogisticRegression

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.config == "c3":
        return bench_sv(args)
    if args.config == "c4":
        return bench_hier(args)
    if args.config == "ecs":
        return bench_ecs(args)
    if args.x_dtype is None:
        # the configurations as stated: C5 in bf16 (the plain-FP32 form timed
        # beside it), C1 and C2 plain fp32
        args.x_dtype = "bf16" if args.config == "c5" else "fp32"
    X, y, beta, cfg = workloads.make(args.config, N=args.rows, chains=args.chains)
    N, D, C = cfg["N"], cfg["D"], cfg["chains"]
    log(f"workload {args.config}: N={N} D={D} chains={C} x_dtype={args.x_dtype}")
    keys = nrandom.chain_init_keys(nrandom.PRNGKey(0), C)
    t0 = time.perf_counter()
    z0, imm, mode, lsd = workloads.laplace_init(X, y, C, full_cov=args.dense_mass,
                                                 return_mode=True)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    imd = torch.diagonal(imm) if imm.ndim == 2 else imm
    log(f"laplace init: {setup_s:.1f}s, inverse mass diagonal in "
        f"[{float(imd.min()):.3g}, {float(imd.max()):.3g}]")

    model = LogisticRegression(
        X, y, x_dtype=args.x_dtype,
        kernel=args.bf16_kernel if args.x_dtype == "bf16" else "auto")
    nw = max(args.warmup, MIN_ADAPT)
    adapt_mass = args.mass != "laplace"
    regularize = args.mass != "adapt-unregularized"
    eng, out, wall, warm_s, kms, kn = run_form(model, C, nw, args.steps, z0, keys,
                                               imm=imm, heuristic=True, dense=args.dense_mass,
                                               adapt_mass=adapt_mass, regularize=regularize,
                                               launch_shape=args.launch_shape)
    ge = eng.grad_evals
    ns = out["num_steps"].cpu().numpy()
    div = int(out["diverging"].sum())
    acc = float(out["accept_prob"].float().mean())
    mess, rhat = min_ess(out["z"])
    gps = ge / wall
    valid = not degenerate(div, acc, C * args.steps)
    ss = eng.get_state().adapt_state.step_size.double().cpu().numpy()
    leaves = ns.sum(1).astype(np.float64)
    spread = {"step_size_q": [float(np.quantile(ss, q)) for q in (0.0, 0.01, 0.5, 0.99, 1.0)],
              "chain_leaves_q": [float(np.quantile(leaves, q)) for q in (0.0, 0.01, 0.5, 0.99, 1.0)],
              "slowest_over_mean_chain": float(leaves.max() / max(leaves.mean(), 1e-9))}
    log(f"per-chain spread: {spread}")
    # posterior mean of the chains against the Laplace mode of the fp32 design,
    # in Laplace standard deviations: how far this storage form moves the
    # posterior (X rounded to bf16 / e4m3 defines a slightly different model)
    shift = ((out["z"].double().mean((0, 1)) - mode) / lsd).abs()
    spread["mean_shift_vs_fp32_laplace_mode_sd"] = {"max": float(shift.max()),
                                                    "rms": float(shift.pow(2).mean().sqrt())}
    log(f"{args.x_dtype}: warmup {warm_s:.1f}s, timed {wall:.2f}s, grad evals {ge}, "
        f"{gps:.4g}/s, min ESS {mess:.1f}, max R-hat {rhat:.3f}, mean tree {ns.mean():.1f}, "
        f"accept {acc:.3f}, divergences {div}, step size "
        f"{float(eng.get_state().adapt_state.step_size.mean()):.3g}")
    # Every tick evaluates only the chains that asked for a gradient (the
    # compacted request list), so the algorithmic work of the timed launches
    # is 4 N D per evaluation times the timed gradient evaluations.
    k_avg_ms = kms / max(kn, 1)
    flops = model.algorithmic_flops(1) * ge / max(kn, 1)
    achieved = model.algorithmic_flops(1) * ge / (kms * 1e-3) / 1e12
    # the fp8 storage form decodes X to fp32 and runs on the VALU
    peak = PEAKS["fp32" if args.x_dtype == "fp8" else args.x_dtype]
    post_state = eng.get_state()

    fp32 = None
    if args.fp32_steps > 0 and args.x_dtype != "fp32" and valid:
        m32 = LogisticRegression(X, y, x_dtype="fp32", kernel="mfma")
        e32, o32, w32, _, k32, n32 = run_form(m32, C, nw, args.fp32_steps, z0, keys,
                                              state=post_state, dense=args.dense_mass,
                                              adapt_mass=adapt_mass, regularize=regularize)
        me32, rh32 = min_ess(o32["z"]) if args.fp32_steps >= 4 else (float("nan"), float("nan"))
        # the bf16 chains' ESS over the same number of draws, for comparison
        mes_b, _ = min_ess(out["z"][:, :args.fp32_steps])
        fp32 = {"grad_evals_per_s": e32.grad_evals / w32, "min_ess_per_s": me32 / w32,
                "min_ess_per_iteration": me32 / args.fp32_steps,
                "bf16_min_ess_per_iteration_same_draws": mes_b / args.fp32_steps,
                "max_rhat": rh32,
                "steps": args.fp32_steps, "kernel_avg_ms": k32 / max(n32, 1),
                "mean_tree": float(o32["num_steps"].float().mean())}
        log(f"fp32 form: {fp32}")
        del e32, m32

    cpu = None
    if args.cpu_budget > 0:
        threads = cpu_threads()
        try:
            cps, cc, calls, cw = cpu_baseline(X, y, z0, args.cpu_budget, threads)
            ess_per_eval = mess / max(ge, 1)
            cpu = {"value": cps, "unit": "grad evals/s", "cores": threads, "kind": "port",
                   "cpu": cpu_model(), "chains": cc,
                   "sample": f"batched potential + gradient of {cc} chains (the vectorized "
                             f"chain method's per-leapfrog work: one pass over the {N}x{D} "
                             f"fp32 design for all chains, {threads} threads), {calls} calls "
                             f"in {cw:.1f}s; tree bookkeeping not included, so this bounds "
                             f"the CPU sampler's gradient rate from above",
                   "min_ess_per_s_at_gpu_ess_per_grad_eval": cps * ess_per_eval,
                   "gpu_over_cpu_grad_evals": gps / cps}
            log(f"cpu baseline: {cpu}")
        except Exception as e:  # the CPU leg is reported, not required
            log(f"cpu baseline skipped: {e!r}")

    line = {
        "metric": ("min-ESS/sec + grad evals/sec, 1024-chain logistic NUTS, MFMA vs plain FP32"
                   if args.config == "c5" else
                   f"min-ESS/sec + grad evals/sec, {C}-chain logistic NUTS ({args.config})"),
        "value": gps,
        "unit": "grad evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": nw,
        "warmup_requested": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.x_dtype == "fp8" else args.x_dtype,
        "data": "synthetic",
        "config": {"workload": f"logistic NUTS {args.config}: N={N}, D={D}, chains={C}, "
                               f"X {args.x_dtype}, max_tree_depth 10, target_accept 0.8, "
                               f"{'dense' if args.dense_mass else 'diag'} mass",
                   "global_batch": C, "seq_len": N,
                   "parallelism": "chains",
                   "init_params": "Laplace mode + Laplace sd * uniform(-2, 2) per chain",
                   "inverse_mass_matrix": ("Laplace covariance (dense_mass=True)"
                                           if args.dense_mass else "Laplace variances (diagonal)"),
                   "adapt_mass_matrix": adapt_mass,
                   "regularize_mass_matrix": regularize,
                   "find_heuristic_step_size": True},
        "valid": valid,
        "setup_s": setup_s,
        "tuning": eng.tuning(),
        "min_ess_per_s": mess / wall,
        "min_ess": mess,
        "min_ess_per_iteration": mess / args.steps,
        "max_rhat": rhat,
        "mean_tree_size": float(ns.mean()),
        "chains": spread,
        "mean_accept_prob": acc,
        "divergences": div,
        "warmup_s": warm_s,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic_bytes(args, N, D, C),
                     "kernel": model_kernel_name(args.x_dtype, N, D, C, args.bf16_kernel == "single"), "kernel_avg_ms": k_avg_ms,
                     "launches": kn, "algorithmic_flops_per_launch": flops,
                     "traffic_per": "launch over all C chains (profile_kernel); a tick's "
                                    "launch covers only the chains that asked for a gradient"},
        "fp32_form": fp32,
        "bf16_vs_fp32_min_ess_per_s": (mess / wall) / fp32["min_ess_per_s"]
        if fp32 and fp32["min_ess_per_s"] > 0 else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    if not valid:
        log(f"degenerate run: {div} of {C * args.steps} timed transitions diverged, mean "
            f"accept {acc:.3f}")
        sys.exit(3)


def bench_sv(args):
    """Stochastic volatility, T=3000, 128 chains: one persistent kernel launch
    per 50 iterations (one wavefront per chain runs NUTS and the adjoint)."""
    import torch
    from numpyro_amd import random as nrandom
    from numpyro_amd.models import StochasticVolatility
    from numpyro_amd.workloads import make_sv
    T = args.rows or 3000
    C = args.chains or 128
    r, info = make_sv(T, 3)
    model = StochasticVolatility(r, noncentred=args.noncentred)
    keys = nrandom.chain_init_keys(nrandom.PRNGKey(0), C)
    z0 = torch.zeros((C, T + 2), dtype=torch.float32, device="cuda")
    z0[:, 0] = float(np.log(5.0))
    z0[:, 1 if args.noncentred else T + 1] = float(np.log(0.02))
    eng, out, wall, warm_s, kms, kn = run_form(model, C, args.warmup, args.steps, z0, keys,
                                               launch_shape=args.launch_shape)
    mess, rhat = min_ess(out["z"])
    ge = eng.grad_evals
    line = {"metric": "min-ESS/sec + grad evals/sec, stochastic volatility NUTS",
            "value": ge / wall, "unit": "grad evals/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
            "data": "synthetic",
            "config": {"workload": f"SV T={T}, chains={C}, "
                                   f"{'non-centred' if args.noncentred else 'centred'}",
                       "global_batch": C, "seq_len": T, "parallelism": "chains"},
            "min_ess_per_s": mess / wall, "min_ess": mess, "max_rhat": rhat,
            "mean_tree_size": float(out["num_steps"].float().mean()),
            "straggler_ratio_50": straggler_ratio(out["num_steps"].cpu().numpy()),
            "mean_accept_prob": float(out["accept_prob"].float().mean()),
            "divergences": int(out["diverging"].sum()), "warmup_s": warm_s,
            "persistent_launch_avg_ms": kms / max(kn, 1), "launches": kn,
            "tuning": eng.tuning(),
            "hbm_bytes_per_grad_eval_per_chain": model.algorithmic_bytes(1)}
    log(json.dumps(line))
    print(json.dumps(line), flush=True)


def bench_hier(args):
    """1024 ragged hierarchical logistic regressions x 4 chains in one
    persistent launch per 50 iterations (a workgroup per model)."""
    import torch
    from numpyro_amd import random as nrandom
    from numpyro_amd.models import HierarchicalLogistic
    from numpyro_amd.workloads import make_hier
    G = args.chains or 1024
    Xs, ys = make_hier(G, 4)
    model = HierarchicalLogistic(Xs, ys, chains_per_group=4)
    C = G * 4
    keys = nrandom.chain_init_keys(nrandom.PRNGKey(0), C)
    z0 = torch.zeros((C, model.dim), dtype=torch.float32, device="cuda")
    eng, out, wall, warm_s, kms, kn = run_form(model, C, args.warmup, args.steps, z0, keys,
                                               launch_shape=args.launch_shape)
    per = group_diagnostics(out["z"], model.dims, 4)
    mess, rhat = per["min_ess"], per["max_rhat"]
    ge = eng.grad_evals
    line = {"metric": "min-ESS/sec + grad evals/sec, ragged hierarchical logistic NUTS",
            "value": ge / wall, "unit": "grad evals/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
            "data": "synthetic",
            "config": {"workload": f"{G} hierarchical logistic regressions, N_i U[50,5000], "
                                   f"D_i U[4,64], 4 chains each", "global_batch": C,
                       "seq_len": int(sum(x.shape[0] for x in Xs)), "parallelism": "chains"},
            "min_ess_per_s": mess / wall, "min_ess": mess, "max_rhat": rhat,
            "diagnostics": "per model over its own 4 chains and D_i + 2 coordinates; "
                           "min ESS / max split R-hat over the 1024 models",
            "groups_rhat_above_1.01": per["n_bad"], "worst_groups": per["worst"],
            "mean_tree_size": float(out["num_steps"].float().mean()),
            "straggler_ratio_50": straggler_ratio(out["num_steps"].cpu().numpy()),
            "mean_accept_prob": float(out["accept_prob"].float().mean()),
            "divergences": int(out["diverging"].sum()), "warmup_s": warm_s,
            "persistent_launch_avg_ms": kms / max(kn, 1), "launches": kn,
            "tuning": eng.tuning(),
            "algorithmic_bytes_per_sweep": model.algorithmic_bytes(C),
            "hbm_GBps_at_grad_rate": model.algorithmic_bytes(C) * ge / C / wall / 1e9}
    log(json.dumps(line))
    print(json.dumps(line), flush=True)
    if args.per_model:
        # per model: split R-hat / ESS over its own chains, and per chain the
        # mean accept probability and the mean of the log-tau coordinate
        acc = out["accept_prob"].float().mean(1).reshape(G, 4).cpu().numpy()
        zc = out["z"].float().mean(1).cpu().numpy()
        lt = np.array([zc[c, model.dims[c // 4] + 1] for c in range(C)]).reshape(G, 4)
        rows = [{"group": g, "N": int(Xs[g].shape[0]), "D": int(model.dims[g]),
                 "max_rhat": float(per["rhat_g"][g]), "min_ess": float(per["ess_g"][g]),
                 "accept": acc[g].round(4).tolist(), "logtau_mean": lt[g].round(4).tolist()}
                for g in range(G)]
        with open(args.per_model, "w") as f:
            json.dump(rows, f)


def straggler_ratio(num_steps, chunk=50):
    """Mean over chunks of `chunk` iterations of (leaves of the slowest
    chain in the chunk) / (mean leaves per chain in the chunk): the factor
    by which a launch that waits for its slowest chain exceeds the mean
    chain's work."""
    ns = np.asarray(num_steps, dtype=np.float64)
    r = []
    for i in range(0, ns.shape[1] - chunk + 1, chunk):
        tot = ns[:, i:i + chunk].sum(1)
        r.append(tot.max() / max(tot.mean(), 1e-9))
    return float(np.mean(r)) if r else float("nan")


def group_diagnostics(z, dims, K):
    """ESS and split R-hat of each model over its own K chains: draws z
    [G*K, S, Dmax] are regrouped to [K, S, G*Dmax] so that every (model,
    coordinate) pair is one parameter of the diagnostics kernel; entries past
    a model's D_i + 2 coordinates (zero padding) are dropped."""
    import torch
    from numpyro_amd.diagnostics import device_diagnostics
    G = len(dims)
    CK, S, Dm = z.shape
    zz = z.reshape(G, K, S, Dm).permute(1, 2, 0, 3).reshape(K, S, G * Dm).contiguous()
    ess, rhat = device_diagnostics(zz)
    ess = ess.reshape(G, Dm).cpu().numpy()
    rhat = rhat.reshape(G, Dm).cpu().numpy()
    mask = np.arange(Dm)[None, :] < (np.asarray(dims)[:, None] + 2)
    ess_g = np.where(mask, ess, np.inf).min(1)
    rhat_g = np.where(mask, rhat, -np.inf).max(1)
    order = np.argsort(-rhat_g)[:5]
    return {"min_ess": float(ess_g.min()), "max_rhat": float(rhat_g.max()),
            "n_bad": int((rhat_g > 1.01).sum()), "rhat_g": rhat_g, "ess_g": ess_g,
            "worst": [{"group": int(g), "D": int(dims[g]), "max_rhat": float(rhat_g[g]),
                       "min_ess": float(ess_g[g])} for g in order]}


def bench_ecs(args):
    """HMCECS on the covtype-sized logistic design (examples/covtype.py:154-165:
    subsample_size 1000, num_blocks 100, Taylor proxy of degree 2): the C5
    data (N=2M, D=256, fp32 X) with 1024 chains, each on its own subsample.
    The proxy's reference point is the data-generating coefficient vector
    (the example uses a MAP estimate)."""
    import torch
    from numpyro_amd import _lib
    from numpyro_amd import random as nrandom
    from numpyro_amd import workloads
    from numpyro_amd.models import SubsampledLogisticRegression
    X, y, beta, cfg = workloads.make("c5", N=args.rows, chains=args.chains)
    N, D = X.shape
    C = args.chains or 1024
    m, nb = args.subsample, args.num_blocks
    t0 = time.perf_counter()
    model = SubsampledLogisticRegression(X, y, m, num_blocks=nb, proxy_degree=2,
                                         reference=torch.as_tensor(beta))
    keys = nrandom.chain_init_keys(nrandom.PRNGKey(0), C)
    kd = torch.as_tensor(np.asarray(keys, np.uint32).view(np.int32)).cuda().contiguous()
    model.keys = kd.clone()
    model.idx = torch.empty((C, m), dtype=torch.int32, device="cuda")
    _lib.check(_lib.load().nax_randint(_lib.ptr(kd), C, m, 0, N, int(nrandom.PARTITIONABLE),
                                       _lib.ptr(model.idx)), "randint")
    # init_to_value(values=ref_params), as the example's inner kernel
    z0 = beta.to(torch.float32).reshape(1, D).repeat(C, 1).contiguous()
    eng, out, wall, warm_s, kms, kn = run_form(model, C, args.warmup, args.steps, z0, keys)
    setup_s = time.perf_counter() - t0 - wall - warm_s
    mess, rhat = min_ess(out["z"])
    ge = eng.grad_evals
    per_eval = model.algorithmic_bytes(1)
    # rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (KiB per dispatch) on the full
    # 1024-chain launch of this workload, scaled by the copy calibration
    traffic = ((2 * 587_350 + 3_105) * 1024
               if (N, D, C, m) == (2_000_000, 256, 1024, 1000) else None)
    ach = ge * per_eval / (kms / 1e3) / 1e9 if kms > 0 else None
    peak = 8000.0
    line = {"metric": "min-ESS/sec + grad evals/sec, 1024-chain HMCECS logistic NUTS",
            "value": ge / wall, "unit": "grad evals/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
            "data": "synthetic",
            "config": {"workload": f"HMCECS logistic NUTS: N={N}, D={D}, chains={C}, "
                                   f"subsample {m}, {nb} blocks, Taylor proxy degree 2",
                       "global_batch": C, "seq_len": m, "parallelism": "chains"},
            "min_ess_per_s": mess / wall, "min_ess": mess, "max_rhat": rhat,
            "mean_tree_size": float(out["num_steps"].float().mean()),
            "straggler_ratio_50": straggler_ratio(out["num_steps"].cpu().numpy()),
            "mean_accept_prob": float(out["accept_prob"].float().mean()),
            "divergences": int(out["diverging"].sum()),
            "gibbs_accept_prob_mean": float(eng.gibbs_accept.mean()),
            "warmup_s": warm_s, "setup_s": setup_s,
            "roofline": {"bound": "hbm", "achieved": ach, "peak": peak,
                         "unit": "GB/s", "frac": ach / peak if ach else None,
                         "traffic": traffic,
                         "traffic_per": "launch over all 1024 chains (profile_kernel --config "
                                        "ecs, FETCH_SIZE x2.0 + WRITE_SIZE x1.0)",
                         "kernel": "ecs_eval_kernel", "kernel_avg_ms": kms / max(kn, 1),
                         "launches": kn, "algorithmic_bytes_per_chain_eval": per_eval}}
    log(json.dumps(line))
    print(json.dumps(line), flush=True)


def model_kernel_name(x_dtype, N, D, C, single=False):
    """The kernel the library's automatic choice runs for this workload."""
    dp = (D + 31) // 32 * 32
    if x_dtype == "bf16":
        return f"logit_bf16_pipe<{dp}, {'true' if single else 'false'}>"
    if x_dtype == "fp8":
        return "logit_partial_f32<XFP8> (VALU tile, e4m3 decoded to fp32)"
    if C <= 16 and D <= 64 and N * D <= (1 << 20):
        return (f"persistent_kernel<LogisticChainModel<{32 if D <= 32 else 64}>> "
                "(one wavefront per chain; a launch runs 50 NUTS iterations)")
    if C < 32:
        return "logit_partial_f32 (VALU tile)"
    if C > 128:
        return f"logit_mfma<FormF32,{dp}>"
    return f"logit_partial_f32 or logit_mfma<FormF32,{dp}> (autotuned)"


# HBM bytes per launch of the main kernel measured with rocprofv3 --pmc
# FETCH_SIZE / WRITE_SIZE (scaled by a 1 GiB copy: x2.0 / x1.0), see
# profiles/logit_bf16_pipe_c5.md.  Only known for the measured workload.
_MEASURED_TRAFFIC = {("single", 2_000_000, 256, 1024): (2 * 509_516 + 33_024) * 1024,
                     ("two-copy", 2_000_000, 256, 1024): (2_016_666 + 33_024) * 1024}


def traffic_bytes(args, N, D, C):
    form = args.bf16_kernel if args.x_dtype == "bf16" else "fp32"
    return _MEASURED_TRAFFIC.get((form, N, D, C))


if __name__ == "__main__":
    main()
