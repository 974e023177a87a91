"""Oracle NUTS on chosen models of the ragged hierarchical workload -- TEST
INFRASTRUCTURE ONLY (a checker run beside the GPU sampler, never the
sampler).

Each model g of make_hier(1024, 4) is sampled by the float64 restatement
(oracle.hmc) with the 4 keys its chains get inside the 4096-chain batch
(chain_keys(PRNGKey(0), 4096)[4g:4g+4], the keys ``bench.py --config c4``
uses), from z = 0, 500 warmup + 500 draws, and its split R-hat / ESS are
computed over its own 4 chains and D_g + 2 coordinates, as the bench does.
With ``--gpu-json`` (the file ``bench.py --config c4 --per-model`` writes)
the GPU's numbers for the same models are printed beside the oracle's.

    python -m oracle.c4_groups 28 71 593 --gpu-json gpurun_out/c4_models.json
"""
import argparse
import json
import time
from multiprocessing import Pool

import numpy as np

from . import hmc as ohmc
from . import jax_random as ojr
from . import models as om
from .diagnostics import effective_sample_size, split_gelman_rubin

_DATA = None


def _data():
    global _DATA
    if _DATA is None:
        from numpyro_amd.workloads import make_hier
        _DATA = make_hier(1024, 4)
    return _DATA


def run_model(g, num_warmup=500, num_samples=500):
    Xs, ys = _data()
    X, y = Xs[g].astype(np.float64), ys[g].astype(np.float64)
    D = X.shape[1]
    keys = np.asarray(ohmc.chain_keys(ojr.PRNGKey(0), 4096))[4 * g:4 * g + 4]
    f = lambda q: om.hier_potential_grad(q, X, y)
    draws, acc, trees = [], [], []
    t0 = time.time()
    for c in range(4):
        ik, sk = ohmc.hmc(f)
        st = ik(np.zeros(D + 2), num_warmup, rng_key=keys[c])
        zs, a = [], []
        for i in range(num_warmup + num_samples):
            st = sk(st)
            if i >= num_warmup:
                zs.append(st.z.copy())
                a.append(st.accept_prob)
                trees.append(st.num_steps)
        draws.append(np.stack(zs))
        acc.append(float(np.mean(a)))
    z = np.stack(draws)
    return {"group": g, "N": int(X.shape[0]), "D": D,
            "max_rhat": float(np.max(split_gelman_rubin(z))),
            "min_ess": float(np.min(effective_sample_size(z))),
            "accept": acc, "logtau_mean": [float(v) for v in z[:, :, D + 1].mean(1)],
            "mean_tree": float(np.mean(trees)), "seconds": time.time() - t0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("groups", type=int, nargs="+")
    ap.add_argument("--gpu-json", default=None)
    ap.add_argument("--procs", type=int, default=4)
    args = ap.parse_args()
    gpu = {}
    if args.gpu_json:
        gpu = {r["group"]: r for r in json.load(open(args.gpu_json))}
    rows = []
    with Pool(args.procs) as p:
        for r in p.imap_unordered(run_model, args.groups):
            if r["group"] in gpu:
                q = gpu[r["group"]]
                r["gpu"] = {"max_rhat": q["max_rhat"], "min_ess": q["min_ess"],
                            "accept": q["accept"]}
            rows.append(r)
            print(json.dumps(r), flush=True)
    o = np.array([r["max_rhat"] for r in rows])
    line = {"models": len(rows), "oracle_rhat_above": {t: int((o > t).sum()) for t in (1.01, 1.05, 1.1)},
            "oracle_median_rhat": float(np.median(o)),
            "oracle_median_ess": float(np.median([r["min_ess"] for r in rows]))}
    if all("gpu" in r for r in rows):
        gr = np.array([r["gpu"]["max_rhat"] for r in rows])
        line.update({"gpu_rhat_above": {t: int((gr > t).sum()) for t in (1.01, 1.05, 1.1)},
                     "gpu_median_rhat": float(np.median(gr)),
                     "gpu_median_ess": float(np.median([r["gpu"]["min_ess"] for r in rows]))})
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
