"""Oracle NUTS on the stochastic-volatility workload -- TEST INFRASTRUCTURE
ONLY (a checker run beside the GPU sampler, never the sampler).

Runs the float64 restatement (oracle.hmc) on the T=3000 returns of
make_sv() for a few of the chains ``bench.py --config c3`` runs: chain c
gets key chain_keys(PRNGKey(0), 128)[c] and the bench's initial point
(nu = 5, sigma = 0.02, s = 0), with the same warmup / sample counts.
Prints, per chain, the mean tree size, the fraction of trees at the depth
cap (1023 leaves), the mean accept probability and divergences, then the
split R-hat and ESS over the chains run.

    python -m oracle.sv_chains --centred --chains 0 1 2 3
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


def run_chain(a):
    c, centred, num_warmup, num_samples = a
    from numpyro_amd.workloads import make_sv
    r, _ = make_sv()
    r = r.astype(np.float64)
    T = r.shape[0]
    z0 = np.zeros(T + 2)
    if centred:  # [log nu, s_1..s_T, log sigma]
        f = lambda q: om.sv_potential_grad(q, r)
        z0[0], z0[T + 1] = np.log(5.0), np.log(0.02)
    else:        # [log nu, log sigma, z_1..z_T]
        f = lambda q: om.sv_nc_potential_grad(q, r)
        z0[0], z0[1] = np.log(5.0), np.log(0.02)
    key = np.asarray(ohmc.chain_keys(ojr.PRNGKey(0), 128))[c]
    ik, sk = ohmc.hmc(f)
    st = ik(z0, num_warmup, rng_key=key)
    t0 = time.time()
    zs, ns, ap, dv = [], [], [], 0
    for i in range(num_warmup + num_samples):
        st = sk(st)
        if i >= num_warmup:
            zs.append(st.z.copy())
            ns.append(st.num_steps)
            ap.append(st.accept_prob)
            dv += int(st.diverging)
    ns = np.asarray(ns)
    return c, np.stack(zs), {"chain": c, "mean_tree": float(ns.mean()),
                             "frac_depth_cap": float((ns >= 1023).mean()),
                             "accept": float(np.mean(ap)), "divergences": dv,
                             "step_size": float(st.adapt_state.step_size),
                             "seconds": time.time() - t0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--centred", action="store_true")
    ap.add_argument("--chains", type=int, nargs="+", default=[0, 1, 2, 3])
    ap.add_argument("--warmup", type=int, default=600)
    ap.add_argument("--samples", type=int, default=600)
    args = ap.parse_args()
    jobs = [(c, args.centred, args.warmup, args.samples) for c in args.chains]
    zs = []
    with Pool(len(jobs)) as p:
        for c, z, rec in p.imap_unordered(run_chain, jobs):
            zs.append(z)
            print(json.dumps(rec), flush=True)
    z = np.stack(zs)
    print(json.dumps({"form": "centred" if args.centred else "non-centred",
                      "chains": len(zs), "max_rhat": float(np.max(split_gelman_rubin(z))),
                      "min_ess": float(np.min(effective_sample_size(z)))}), flush=True)


if __name__ == "__main__":
    main()
