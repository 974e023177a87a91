"""Whole-chain parity of the GPU engine against the float64 oracle.

With identical per-chain keys, the engine draws the same threefry bits as
the oracle, so every doubling direction, transition and U-turn decision is
the same unless float32 rounding moves a comparison across its threshold.
Without adaptation every decision is far from its threshold on these
problems and whole chains must match (tree sizes exactly, positions to
float32 accuracy).  With adaptation, the early dual-averaging steps amplify
rounding: the float64 oracle run against itself with its potential
evaluated at float32-rounded positions drifts to ~4e-4 after 40 warmup
iterations, enough to flip a near-tie multinomial transition later on.
There the tree sizes must still match for every iteration, positions must
match over the first 25 iterations."""
import numpy as np
import pytest
import torch

from oracle import hmc as ohmc
from oracle import jax_random as ojr
from oracle import models as om

pytestmark = pytest.mark.gpu


def _problem(N, D, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(N, D)).astype(np.float32)
    beta = rng.normal(size=D)
    y = (rng.uniform(size=N) < 1 / (1 + np.exp(-X @ beta))).astype(np.float32)
    return X, y


def _run_engine(X, y, z0, keys, num_warmup, num_samples, algo="NUTS", **kw):
    from numpyro_amd.infer import hmc
    from numpyro_amd.infer.util import fori_collect
    from numpyro_amd.models import LogisticRegression
    m = LogisticRegression(X, y)
    init_kernel, sample_kernel = hmc(m, algo=algo)
    st = init_kernel(torch.from_numpy(z0), num_warmup, rng_key=keys, **kw)
    out = fori_collect(0, num_warmup + num_samples, sample_kernel, st,
                       transform=lambda s: (s.z, s.num_steps, s.potential_energy,
                                            s.accept_prob))
    return [o.cpu().numpy() for o in out]


def _run_oracle(X, y, z0, keys, num_warmup, num_samples, algo="NUTS", **kw):
    f = lambda q: om.logistic_potential_grad(q, X, y, 1.0)  # noqa: E731
    C = z0.shape[0]
    res = {"z": [], "num_steps": [], "potential_energy": [], "accept_prob": []}
    for c in range(C):
        ik, sk = ohmc.hmc(f, algo=algo)
        st = ik(z0[c].astype(np.float64), num_warmup, rng_key=keys[c], **kw)
        rows = {k: [] for k in res}
        for _ in range(num_warmup + num_samples):
            st = sk(st)
            for k in res:
                rows[k].append(np.copy(getattr(st, k)))
        for k in res:
            res[k].append(np.stack(rows[k]))
    return [np.stack(res[k]) for k in ("z", "num_steps", "potential_energy", "accept_prob")]


@pytest.mark.parametrize("seed, adapt", [(11, True), (12, False)])
def test_nuts_chain_matches_oracle(seed, adapt):
    N, D, C = 200, 5, 3
    X, y = _problem(N, D, seed)
    rng = np.random.default_rng(seed + 100)
    z0 = rng.uniform(-2, 2, size=(C, D)).astype(np.float32)
    keys = np.stack([ojr.split(ojr.PRNGKey(seed), C)[c] for c in range(C)])
    kw = dict(adapt_step_size=adapt, adapt_mass_matrix=adapt,
              step_size=1.0 if adapt else 0.2)
    num_warmup, num_samples = 40, 20
    gz, gns, gpe, gap = _run_engine(X, y, z0, keys, num_warmup, num_samples, **kw)
    oz, ons, ope, oap = _run_oracle(X, y, z0, keys, num_warmup, num_samples, **kw)
    assert np.array_equal(gns, ons), (gns, ons)
    horizon = 25 if adapt else num_warmup + num_samples
    np.testing.assert_allclose(gz[:, :horizon], oz[:, :horizon], rtol=1e-3, atol=1e-3)
    np.testing.assert_allclose(gpe[:, :horizon], ope[:, :horizon], rtol=1e-4 if adapt else 1e-5)
    np.testing.assert_allclose(gap[:, :horizon], oap[:, :horizon], rtol=1e-3, atol=1e-4)


def test_hmc_chain_matches_oracle():
    N, D, C = 150, 4, 2
    X, y = _problem(N, D, 21)
    z0 = np.random.default_rng(3).uniform(-2, 2, size=(C, D)).astype(np.float32)
    keys = np.stack([ojr.split(ojr.PRNGKey(5), C)[c] for c in range(C)])
    kw = dict(trajectory_length=1.0, step_size=0.1)
    gz, gns, gpe, gap = _run_engine(X, y, z0, keys, 30, 20, algo="HMC", **kw)
    oz, ons, ope, oap = _run_oracle(X, y, z0, keys, 30, 20, algo="HMC", **kw)
    assert np.array_equal(gns, ons)
    np.testing.assert_allclose(gz, oz, rtol=1e-3, atol=1e-3)
    np.testing.assert_allclose(gpe, ope, rtol=1e-5)
