"""The C restatement (CPU baseline) against the float64 numpy oracle: same
keys, same tree sizes; and its posterior on a small logistic problem."""
import os

import numpy as np
import pytest

from oracle import hmc as ohmc
from oracle import jax_random as ojr
from oracle import models as om

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                   "libcpu_nuts.so")


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(LIB):
        from oracle import build
        build.build()


def _problem(N, D, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(N, D)).astype(np.float32)
    beta = rng.normal(size=D)
    y = (rng.uniform(size=N) < 1 / (1 + np.exp(-X @ beta))).astype(np.float32)
    return X, y


def test_cpu_nuts_tree_sizes_match_numpy_oracle():
    from oracle import cpu_baseline
    X, y = _problem(200, 5, 12)
    C = 2
    z0 = np.random.default_rng(112).uniform(-2, 2, size=(C, 5)).astype(np.float32)
    keys = ojr.split(ojr.PRNGKey(12), C)
    _, ns, _ = cpu_baseline.nuts_logistic(X, y, z0, keys, 0, 30, step_size=0.2, adapt=False,
                                          threads=1)
    f = lambda q: om.logistic_potential_grad(q, X, y, 1.0)  # noqa: E731
    for c in range(C):
        ik, sk = ohmc.hmc(f)
        st = ik(z0[c].astype(np.float64), 0, rng_key=keys[c], step_size=0.2,
                adapt_step_size=False, adapt_mass_matrix=False)
        want = []
        for _ in range(30):
            st = sk(st)
            want.append(st.num_steps)
        assert list(ns[c]) == want


def test_cpu_nuts_posterior():
    from oracle import cpu_baseline
    rng = np.random.default_rng(0)
    X = rng.normal(size=(3000, 3)).astype(np.float32)
    y = (rng.uniform(size=3000) < 1 / (1 + np.exp(-X @ np.array([1.0, 2.0, 3.0])))).astype(np.float32)
    z0 = rng.uniform(-2, 2, size=(4, 3)).astype(np.float32)
    z, ns, ge = cpu_baseline.nuts_logistic(X, y, z0, ojr.split(ojr.PRNGKey(1), 4), 300, 300)
    assert ge == ns.sum()
    np.testing.assert_allclose(z.reshape(-1, 3).mean(0), [1, 2, 3], atol=0.4)
