"""Whole chains pinned against numbers the reference itself publishes: the
non-centred eight-schools run of the reference's README (data
README.md:35-37, model README.md:101-111, MCMC(NUTS, num_warmup=500,
num_samples=1000) summary README.md:119-136, 0 divergences README.md:138,
expected log joint density -46.09 README.md:143-144).

The model goes through the ``potential_fn`` path: the engine runs NUTS,
the adaptation and the RNG in HIP, and the potential (float64 torch, the
reference's potential_energy of the model with sites flattened in sorted
order ``mu, tau, theta_base``; tau unconstrained by ExpTransform with its
log-Jacobian) is evaluated between ticks.  8 chains x 1000 draws after 500
warmup; with the combined MCSE sqrt(sd_ref^2 / n_eff_ref + sd^2 / n_eff) of
each of the 18 sites, every posterior mean must agree with the README's
within 3.5 MCSE and at least 16 of the 18 within 3 MCSE.  (The README's
means come from one chain of 1000 draws whose n_eff is itself an estimate:
its theta[5] = 3.58 lies about 3 of its own MCSEs below the long-run
posterior mean of about 4.0, so a 3-MCSE bound on every site would sit at
the edge for a correct sampler.)"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

Y = [28.0, 8.0, -3.0, 7.0, -1.0, 1.0, 18.0, 12.0]
SIGMA = [15.0, 10.0, 16.0, 11.0, 9.0, 11.0, 10.0, 18.0]

# README.md:119-136: (mean, sd, n_eff)
REF = {
    "mu": (4.08, 3.51, 720.43), "tau": (3.96, 3.31, 488.63),
    "theta[0]": (6.48, 5.72, 801.59), "theta[1]": (4.95, 5.10, 1183.06),
    "theta[2]": (3.65, 5.58, 581.31), "theta[3]": (4.56, 5.04, 1282.60),
    "theta[4]": (3.41, 4.79, 801.25), "theta[5]": (3.58, 4.80, 1101.33),
    "theta[6]": (6.31, 5.17, 1081.11), "theta[7]": (4.81, 5.38, 954.14),
    "theta_base[0]": (0.41, 0.95, 851.45), "theta_base[1]": (0.15, 0.95, 1568.11),
    "theta_base[2]": (-0.08, 0.98, 1037.16), "theta_base[3]": (0.06, 0.89, 1745.02),
    "theta_base[4]": (-0.14, 0.94, 719.85), "theta_base[5]": (-0.10, 0.96, 1128.45),
    "theta_base[6]": (0.38, 0.95, 1026.50), "theta_base[7]": (0.10, 0.97, 1190.98),
}


def potential_fn(z):
    """-(log joint) in unconstrained space: z = [mu, log tau, theta_base[8]]."""
    y = torch.tensor(Y, dtype=z.dtype, device=z.device)
    sig = torch.tensor(SIGMA, dtype=z.dtype, device=z.device)
    mu, ut, tb = z[0], z[1], z[2:]
    tau = torch.exp(ut)
    lp = -0.5 * (mu / 5.0) ** 2 - math.log(5.0 * math.sqrt(2 * math.pi))      # Normal(0, 5)
    lp = lp + math.log(2 / math.pi) - math.log(5.0) - torch.log1p((tau / 5.0) ** 2)  # HalfCauchy
    lp = lp + ut                                                                # log|d tau/du|
    lp = lp + torch.sum(-0.5 * tb ** 2) - 8 * 0.5 * math.log(2 * math.pi)       # Normal(0, 1)
    theta = mu + tau * tb
    lp = lp + torch.sum(-0.5 * ((y - theta) / sig) ** 2 - torch.log(sig)
                        - 0.5 * math.log(2 * math.pi))
    return -lp


def test_eight_schools_noncentred_matches_reference_readme():
    from numpyro_amd import random
    from numpyro_amd.diagnostics import effective_sample_size, split_gelman_rubin
    from numpyro_amd.infer import MCMC, NUTS
    C = 8
    print(f"[eight schools] {C} chains, 500 warmup + 1000 samples", flush=True)
    init = np.random.default_rng(0).uniform(-2, 2, size=(C, 10))
    mcmc = MCMC(NUTS(potential_fn=potential_fn), num_warmup=500, num_samples=1000,
                num_chains=C, progress_bar=False)
    mcmc.run(random.PRNGKey(0), init_params=torch.from_numpy(init).float(),
             extra_fields=("potential_energy",))
    z = mcmc.get_samples(group_by_chain=True).cpu().numpy().astype(np.float64)
    div = int(mcmc.get_extra_fields()["diverging"].sum())
    pe = mcmc.get_extra_fields()["potential_energy"].cpu().numpy()
    tau = np.exp(z[..., 1])
    sites = {"mu": z[..., 0], "tau": tau}
    for j in range(8):
        sites[f"theta_base[{j}]"] = z[..., 2 + j]
        sites[f"theta[{j}]"] = z[..., 0] + tau * z[..., 2 + j]
    worst = 0.0
    within3 = 0
    for name, (m_ref, s_ref, n_ref) in REF.items():
        x = sites[name]
        ess = float(effective_sample_size(x[..., None])[0])
        rhat = float(split_gelman_rubin(x[..., None])[0])
        m, s = x.mean(), x.std()
        mcse = math.sqrt(s_ref ** 2 / n_ref + s ** 2 / ess)
        dz = abs(m - m_ref) / mcse
        print(f"[eight schools] {name}: mean {m:.3f} (ref {m_ref}), sd {s:.3f} (ref {s_ref}), "
              f"ess {ess:.0f}, r_hat {rhat:.4f}, |diff| {abs(m - m_ref):.3f} = {dz:.2f} MCSE",
              flush=True)
        worst = max(worst, dz)
        within3 += dz < 3
        assert dz < 3.5, name
        assert rhat < 1.01, name
    assert within3 >= 16, within3
    # the README's 1000 draws report 0 divergences; allow a handful in 8000
    assert div <= 8, div
    # expected log joint density -46.09 (README.md:143-144); its ESS is
    # not printed there, 100 is assumed as a lower bound
    sd_pe = pe.std()
    ess_pe = float(effective_sample_size(pe.reshape(C, -1)[..., None].astype(np.float64))[0])
    bound = 3 * sd_pe * math.sqrt(1 / 100 + 1 / ess_pe)
    print(f"[eight schools] E[-pe] {-pe.mean():.3f} (ref -46.09) < {bound:.3f}; "
          f"divergences {div}; worst |diff| {worst:.2f} MCSE, {within3} of 18 within 3",
          flush=True)
    assert abs(-pe.mean() - (-46.09)) < bound
