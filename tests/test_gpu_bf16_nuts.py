"""NUTS with the BF16-MFMA potential: 64 chains on a logistic problem, the
posterior mean must agree with a float64-oracle reference within 3 MCSE
(the reference posterior mean/sd come from the fp32 engine run, whose tree
parity with the oracle is pinned in test_gpu_nuts.py)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_bf16_matches_fp32_posterior():
    from numpyro_amd import random
    from numpyro_amd.diagnostics import effective_sample_size, split_gelman_rubin
    from numpyro_amd.infer import MCMC, NUTS
    from numpyro_amd.models import LogisticRegression
    rng = np.random.default_rng(3)
    N, D, C = 5000, 8, 64
    X = rng.normal(size=(N, D)).astype(np.float32)
    beta = rng.normal(size=D) * 0.5
    y = (rng.uniform(size=N) < 1 / (1 + np.exp(-X @ beta))).astype(np.float32)
    res = {}
    for xd in ("fp32", "bf16"):
        m = LogisticRegression(X, y, x_dtype=xd, kernel="mfma" if xd == "fp32" else "auto")
        mcmc = MCMC(NUTS(m), num_warmup=300, num_samples=300, num_chains=C)
        mcmc.run(random.PRNGKey(0), init_params=torch.zeros(C, D))
        z = mcmc.get_samples(group_by_chain=True)["coefs"].cpu().numpy().astype(np.float64)
        assert np.nanmax(split_gelman_rubin(z)) < 1.01
        res[xd] = (z.reshape(-1, D).mean(0), z.reshape(-1, D).std(0), effective_sample_size(z))
    m32, s32, e32 = res["fp32"]
    m16, s16, e16 = res["bf16"]
    mcse = np.sqrt(s32 ** 2 / e32 + s16 ** 2 / e16)
    assert np.all(np.abs(m32 - m16) < 3 * mcse + 1e-3 * s32), (m32 - m16) / mcse
