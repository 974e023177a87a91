"""Pins of the HMCECS oracle (oracle/ecs.py) on the CPU."""
import itertools

import numpy as np
import pytest

from oracle import ecs
from oracle.jax_random import PRNGKey, random_bits, split


def _data(n=60, d=3, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, d))
    beta = np.array([0.8, -0.5, 0.3])[:d]
    y = (rng.uniform(size=n) < 1 / (1 + np.exp(-X @ beta))).astype(np.float64)
    return X, y, beta


@pytest.mark.parametrize("part", [True, False])
@pytest.mark.parametrize("span", [1, 2, 7, 100, 1000, 65535])
def test_randint_is_exact_modulus_for_small_spans(part, span):
    # (hi % s) * (2^32 mod s) + lo % s == (hi * 2^32 + lo) mod s while no
    # uint32 product wraps, i.e. for spans below 2^16.
    key = PRNGKey(17)
    k1, k2 = split(key, 2, part)
    hi = random_bits(k1, 50, part).astype(object)
    lo = random_bits(k2, 50, part).astype(object)
    want = np.array([(int(h) * 2 ** 32 + int(l)) % span for h, l in zip(hi, lo)]) + 3
    got = ecs.randint(key, 50, 3, 3 + span, part)
    np.testing.assert_array_equal(got, want)


def test_randint_range_and_uniformity():
    from scipy.stats import chisquare
    got = ecs.randint(PRNGKey(5), 20000, 0, 10)
    assert got.min() >= 0 and got.max() < 10
    assert chisquare(np.bincount(got, minlength=10)).pvalue > 1e-4
    big = ecs.randint(PRNGKey(6), 20000, 0, 2_000_000)
    assert big.min() >= 0 and big.max() < 2_000_000
    assert abs(big.mean() / 1e6 - 1.0) < 0.02
    assert int(ecs.randint(PRNGKey(1), None, 4, 4)) == 4  # empty span returns minval


@pytest.mark.parametrize("m,nb,bs,pad", [(10, 3, 4, 2), (10, 4, 3, 2), (10, 6, 2, 0),
                                         (1000, 100, 10, 0), (7, 1, 7, 0), (5, 10, 1, 0)])
def test_block_layout(m, nb, bs, pad):
    assert ecs.block_layout(m, nb) == (bs, pad)
    assert (m + pad) % bs == 0


@pytest.mark.parametrize("m,nb", [(10, 3), (10, 6), (12, 4), (9, 1)])
def test_update_block_changes_one_clamped_block(m, nb):
    idx = np.arange(100, 100 + m, dtype=np.int32)
    bs, pad = ecs.block_layout(m, nb)
    for s in range(40):
        new = ecs.update_block(PRNGKey(s), idx, nb, 50)
        changed = np.nonzero(new != idx)[0]
        assert new.shape == (m,) and new.dtype == np.int32
        assert np.all((new[changed] >= 0) & (new[changed] < 50))
        # the redrawn slots form one block: start = min(chosen * bs, m + pad - bs)
        starts = [min(c * bs, m + pad - bs) for c in range(nb)]
        assert any(np.all((changed >= st) & (changed < min(st + bs, m))) for st in starts)
        assert changed.size >= 1


def _fd_grad_hess(f, x, h=1e-4):
    d = x.size
    g = np.zeros(d)
    H = np.zeros((d, d))
    for i in range(d):
        e = np.zeros(d); e[i] = h
        g[i] = (f(x + e) - f(x - e)) / (2 * h)
        for j in range(d):
            e2 = np.zeros(d); e2[j] = h
            H[i, j] = (f(x + e + e2) - f(x + e - e2) - f(x - e + e2) + f(x - e - e2)) / (4 * h * h)
    return g, H


def test_taylor_terms_against_finite_difference_expansion():
    # test/infer/test_hmc_gibbs.py:333-399 with derivatives of the per-row
    # log-likelihood taken numerically instead of by autodiff.
    X, y, beta = _data()
    ref = beta + 0.05
    st = ecs.taylor_stats(X, y, ref)
    pert = ref + np.array([0.1, -0.2, 0.15])
    idx = np.array([3, 3, 17, 40, 59])
    d = pert - ref
    rows = []
    for i in range(X.shape[0]):
        f = lambda b, i=i: ecs.row_loglik(X[i] @ b, y[i])
        g, H = _fd_grad_hess(f, ref)
        rows.append(f(ref) + g @ d + 0.5 * d @ H @ d)
    rows = np.array(rows)
    full, sub = ecs.proxy(pert, idx, X, st)
    np.testing.assert_allclose(sub, rows[idx], rtol=1e-6, atol=1e-8)
    np.testing.assert_allclose(full, rows.sum(), rtol=1e-7)
    st1 = ecs.taylor_stats(X, y, ref, degree=1)
    full1, _ = ecs.proxy(pert, idx, X, st1)
    np.testing.assert_allclose(full1, st["LL"] + st["G"] @ d, rtol=1e-12)


def test_estimator_at_reference_is_full_loglik():
    X, y, beta = _data()
    st = ecs.taylor_stats(X, y, beta)
    full_ll = ecs.row_loglik(X @ beta, y).sum()
    for s in range(3):
        idx = np.random.default_rng(s).integers(0, X.shape[0], 7)
        np.testing.assert_allclose(ecs.loglik_estimate(beta, idx, X, y, X.shape[0], st),
                                   full_ll, rtol=1e-12)


@pytest.mark.parametrize("with_proxy", [True, False])
def test_mean_term_unbiased_over_all_size_one_subsamples(with_proxy):
    # m = 1: the variance term vanishes and averaging over every index gives
    # the full-data log-likelihood exactly.
    X, y, beta = _data(n=25)
    st = ecs.taylor_stats(X, y, beta + 0.3) if with_proxy else None
    b = beta - 0.2
    n = X.shape[0]
    est = [ecs.loglik_estimate(b, np.array([i]), X, y, n, st) for i in range(n)]
    np.testing.assert_allclose(np.mean(est), ecs.row_loglik(X @ b, y).sum(), rtol=1e-12)


def test_variance_term_over_all_pairs():
    # m = 2 with replacement: E[var_ddof0(d)] = sigma^2 / 2 (population
    # variance of the diffs), so the expected correction is n^2 sigma^2 / 8.
    X, y, beta = _data(n=12)
    st = ecs.taylor_stats(X, y, beta + 0.4)
    b = beta - 0.3
    n = X.shape[0]
    full, sub = ecs.proxy(b, np.arange(n), X, st)
    diff = ecs.row_loglik(X @ b, y) - sub
    est = [ecs.loglik_estimate(b, np.array(p), X, y, n, st)
           for p in itertools.product(range(n), repeat=2)]
    want = full + diff.sum() - n * n / 8.0 * diff.var()
    np.testing.assert_allclose(np.mean(est), want, rtol=1e-12)


@pytest.mark.parametrize("with_proxy,degree", [(True, 2), (True, 1), (False, 2)])
def test_gradient_against_central_differences(with_proxy, degree):
    X, y, beta = _data()
    st = ecs.taylor_stats(X, y, beta + 0.1, degree) if with_proxy else None
    idx = np.array([1, 5, 5, 22, 31, 47, 58])
    b = beta + np.array([0.2, 0.1, -0.3])
    u, g = ecs.potential_grad(b, idx, X, y, X.shape[0], st, 1.5)
    h = 1e-6
    fd = [(ecs.potential(b + h * e, idx, X, y, X.shape[0], st, 1.5)
           - ecs.potential(b - h * e, idx, X, y, X.shape[0], st, 1.5)) / (2 * h)
          for e in np.eye(3)]
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-6)


def test_naive_potential_with_every_row_is_full_potential():
    from oracle.models import logistic_potential
    X, y, beta = _data()
    idx = np.arange(X.shape[0])
    np.testing.assert_allclose(ecs.potential(beta, idx, X, y, X.shape[0]),
                               logistic_potential(beta, X, y), rtol=1e-12)


def test_planted_sign_error_in_hessian_fails_the_expansion_pin():
    X, y, beta = _data()
    st = ecs.taylor_stats(X, y, beta)
    st["w"] = -st["w"]
    pert = beta + np.array([0.3, -0.2, 0.1])
    d = pert - beta
    f = lambda b: ecs.row_loglik(X[4] @ b, y[4])
    g, H = _fd_grad_hess(f, beta)
    _, sub = ecs.proxy(pert, np.array([4]), X, st)
    assert abs(sub[0] - (f(beta) + g @ d + 0.5 * d @ H @ d)) > 1e-4


def test_gibbs_step_accepts_by_the_uniform():
    X, y, beta = _data()
    st = ecs.taylor_stats(X, y, beta)
    idx = np.random.default_rng(0).integers(0, X.shape[0], 10).astype(np.int32)
    pe, _ = ecs.potential_grad(beta, idx, X, y, X.shape[0], st)
    k, idx2, pe2, g2, ap = ecs.gibbs_step(PRNGKey(3), idx, beta, pe, X, y, X.shape[0], 5, st)
    # at the reference point every subsample gives the same estimate
    assert ap == 1.0 and g2 is not None
    np.testing.assert_allclose(pe2, pe, rtol=1e-12)
    np.testing.assert_array_equal(k, split(PRNGKey(3))[0])


def test_gibbs_step_rejects_a_nan_proposal():
    # hmc_gibbs.py:660-661: clip(exp(pe - pe_new), None, 1) is NaN for a NaN
    # proposal energy, and bernoulli(key, NaN) = uniform < NaN is false
    import warnings
    rng = np.random.default_rng(0)
    N, D, m = 50, 3, 10
    X = rng.normal(size=(N, D))
    X[:25, 0] = np.nan
    y = (rng.uniform(size=N) < 0.5).astype(np.float64)
    idx = rng.integers(25, N, m)
    beta = np.zeros(D)
    pe = ecs.potential(beta, idx, X, y, N)
    assert np.isfinite(pe)
    seen = 0
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        for s in range(20):
            key = np.array([0, s], np.uint32)
            k, idx_new, pe_new, g, ap = ecs.gibbs_step(key, idx, beta, pe, X, y, N, 1)
            if np.isnan(ap):
                seen += 1
                assert g is None and pe_new == pe
                np.testing.assert_array_equal(idx_new, idx)
    assert seen >= 10
