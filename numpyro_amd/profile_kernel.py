"""Times the logistic potential+gradient kernels on a benchmark workload.

    python -m numpyro_amd.profile_kernel --config c5 --x-dtype bf16 --reps 20
    python -m numpyro_amd.profile_kernel --copy-mib 1024     # copy-kernel calibration
    python -m numpyro_amd.profile_kernel --config c4         # ragged hierarchical models
    python -m numpyro_amd.profile_kernel --config ecs        # HMCECS subsampled potential

Each rep evaluates U and dU for every chain at fixed positions (one launch
of the main kernel plus the fixed-order reduction), timed with HIP events
on the sampler's stream.  ``--copy-mib`` also runs a plain device copy of
that many MiB (torch's copy kernel) so that hardware byte counters read on
the main kernel can be scaled by what they read for a copy of known size.
"""
import argparse
import ctypes as C
import json
import sys

import torch


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="c5", choices=["c5", "c2", "c1", "c4", "c3", "c3nc", "ecs"])
    ap.add_argument("--x-dtype", default="bf16", choices=["bf16", "fp32"])
    ap.add_argument("--kernel", default="auto", choices=["auto", "tile", "mfma", "single", "two-copy"])
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--rows", type=int, default=None)
    ap.add_argument("--chains", type=int, default=None)
    ap.add_argument("--copy-mib", type=int, default=0)
    ap.add_argument("--launch-shape", type=int, default=None,
                    help="c4: chains per workgroup + 32 x (waves per chain - 1); "
                         "c1, c3, c3nc: waves per chain")
    args = ap.parse_args(argv)

    from . import _lib, workloads
    from .infer.hmc import Engine
    from .models import LogisticRegression

    out = {}
    if args.reps > 0:
        if args.config in ("c3", "c3nc"):
            from .models import StochasticVolatility
            r, meta = workloads.make_sv(3000, 3)
            model = StochasticVolatility(r, noncentred=args.config == "c3nc")
            C_, D = args.chains or 128, model.dim
            cfg = {"N": 3000, "chains": C_, "D": D}
            gen = torch.Generator(device="cuda")
            gen.manual_seed(0)
            z = 0.1 * torch.randn((C_, D), generator=gen, device="cuda")
            args.x_dtype = "fp32"
        elif args.config == "c4":
            from .models import HierarchicalLogistic
            G = args.chains or 1024
            model = HierarchicalLogistic(*workloads.make_hier(G, 4), chains_per_group=4)
            C_, D = 4 * G, model.dim
            cfg = {"N": sum(model.rows), "chains": C_, "D": D}
            gen = torch.Generator(device="cuda")
            gen.manual_seed(0)
            z = 0.5 * torch.randn((C_, D), generator=gen, device="cuda")
            args.x_dtype = "fp32"
        elif args.config == "ecs":
            # the N=2M, D=256 design, 1024 chains, each on its own subsample of
            # 1000 rows, Taylor proxy of degree 2 around the generating beta
            import numpy as np
            from . import random as nrandom
            from .models import SubsampledLogisticRegression
            X, y, beta, cfg = workloads.make("c5", N=args.rows, chains=args.chains)
            C_, D = cfg["chains"], cfg["D"]
            model = SubsampledLogisticRegression(X, y, 1000, proxy_degree=2, reference=beta)
            keys = nrandom.chain_init_keys(nrandom.PRNGKey(0), C_)
            kd = torch.as_tensor(np.asarray(keys, np.uint32).view(np.int32)).cuda()
            model.keys = kd.contiguous()
            model.idx = torch.empty((C_, 1000), dtype=torch.int32, device="cuda")
            _lib.check(_lib.load().nax_randint(_lib.ptr(model.keys), C_, 1000, 0, X.shape[0],
                                               int(nrandom.PARTITIONABLE), _lib.ptr(model.idx)),
                       "randint")
            z = (beta[None, :] + 1e-3 * torch.randn((C_, D), device="cuda")).contiguous()
            args.x_dtype = "fp32"
        else:
            X, y, beta, cfg = workloads.make(args.config, N=args.rows, chains=args.chains)
            C_, D = cfg["chains"], cfg["D"]
            model = LogisticRegression(X, y, x_dtype=args.x_dtype, kernel=args.kernel)
            z = (beta[None, :] + 1e-3 * torch.randn((C_, D), device="cuda")).contiguous()
        eng = Engine(model, C_, "NUTS", 0)
        if args.launch_shape:
            eng.set_launch_shape(args.launch_shape)
        pe = torch.empty(C_, dtype=torch.float64, device="cuda")
        g = torch.empty((C_, D), dtype=torch.float32, device="cuda")
        lib = _lib.load()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        times = []
        for _ in range(args.reps):
            torch.cuda.synchronize()
            ev[0].record()
            _lib.check(lib.nax_potential_grad(eng.h, _lib.ptr(z), D, C_, _lib.ptr(pe),
                                              _lib.ptr(g)), "potential")
            ev[1].record()
            torch.cuda.synchronize()
            times.append(ev[0].elapsed_time(ev[1]))
        flops = model.algorithmic_flops(C_)
        best = min(times)
        out.update({"config": args.config, "x_dtype": args.x_dtype, "N": cfg["N"], "D": D,
                    "chains": C_, "eval_ms_min": best,
                    "eval_ms_median": sorted(times)[len(times) // 2],
                    "algorithmic_tflops_at_min": flops / (best * 1e-3) / 1e12,
                    "algorithmic_GBps_at_min": model.algorithmic_bytes(C_) / (best * 1e-3) / 1e9})
        print(f"[profile] {args.x_dtype} eval: min {best:.3f} ms, median "
              f"{out['eval_ms_median']:.3f} ms", file=sys.stderr, flush=True)
    if args.copy_mib > 0:
        n = args.copy_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        out["copy_mib"] = args.copy_mib
        print(f"[profile] copied {args.copy_mib} MiB three times", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
