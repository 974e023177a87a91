"""Pins the oracle's log-densities to scipy.stats (as test/test_distributions.py
does) and its analytic gradients to central finite differences."""
import numpy as np
from numpy.testing import assert_allclose
import pytest
import scipy.stats as st

from oracle import distributions as d
from oracle import models as m


def test_log_probs_vs_scipy():
    rng = np.random.RandomState(0)
    x = rng.normal(size=20)
    xp = np.abs(x) + 0.1
    assert_allclose(d.normal_log_prob(x, 0.3, 1.7), st.norm(0.3, 1.7).logpdf(x), atol=1e-12)
    assert_allclose(d.student_t_log_prob(x, 3.3, 0.2, 1.4), st.t(3.3, 0.2, 1.4).logpdf(x),
                    atol=1e-12)
    assert_allclose(d.cauchy_log_prob(x, 0.1, 2.0), st.cauchy(0.1, 2.0).logpdf(x), atol=1e-12)
    assert_allclose(d.half_cauchy_log_prob(xp, 1.5), st.halfcauchy(0, 1.5).logpdf(xp),
                    atol=1e-12)
    assert_allclose(d.exponential_log_prob(xp, 50.0), st.expon(scale=1 / 50).logpdf(xp),
                    atol=1e-10)
    y = (rng.uniform(size=20) < 0.5).astype(float)
    p = 1 / (1 + np.exp(-x))
    assert_allclose(d.bernoulli_logits_log_prob(x, y), st.bernoulli(p).logpmf(y), atol=1e-12)
    # GRW: increments are N(0, scale), start N(0, scale)
    s = rng.normal(size=7)
    want = st.norm(0, 0.4).logpdf(s[0]) + st.norm(0, 0.4).logpdf(np.diff(s)).sum()
    assert_allclose(d.gaussian_random_walk_log_prob(s, 0.4), want, atol=1e-12)


@pytest.mark.parametrize("tf", [d.exp_transform, d.softplus_transform])
def test_transform_log_det(tf):
    x = np.linspace(-3, 3, 11)
    y, ladj = tf(x)
    h = 1e-6
    dydx = (tf(x + h)[0] - tf(x - h)[0]) / (2 * h)
    assert_allclose(ladj, np.log(np.abs(dydx)), atol=1e-6)
    if tf is d.softplus_transform:
        assert_allclose(d.softplus_inv(y), x, atol=1e-10)


def test_logistic_grad_fd():
    rng = np.random.RandomState(1)
    X = rng.normal(size=(50, 4))
    y = (rng.uniform(size=50) < 0.5).astype(float)
    b = rng.normal(size=4)
    u, g = m.logistic_potential_grad(b, X, y, 1.3)
    fd = m.central_fd(lambda z: m.logistic_potential(z, X, y, 1.3), b)
    assert_allclose(g, fd, rtol=1e-6, atol=1e-6)
    # U against a direct log-prob sum (scipy pinned)
    want = -(st.bernoulli(1 / (1 + np.exp(-X @ b))).logpmf(y).sum()
             + st.norm(0, 1.3).logpdf(b).sum())
    assert_allclose(u, want, rtol=1e-12)


def test_sv_grad_fd():
    rng = np.random.RandomState(2)
    T = 12
    r = rng.standard_t(4, size=T) * 0.02
    z = np.concatenate([[np.log(5.0)], rng.normal(size=T) * 0.3 - 4.0, [np.log(0.05)]])
    u, g = m.sv_potential_grad(z, r)
    fd = m.central_fd(lambda zz: m.sv_potential(zz, r), z)
    assert_allclose(g, fd, rtol=1e-5, atol=1e-5)
    nu, sig, s = np.exp(z[0]), np.exp(z[-1]), z[1:-1]
    want = (st.expon(scale=1 / 50).logpdf(sig) + z[-1] + st.expon(scale=10).logpdf(nu) + z[0]
            + st.norm(0, sig).logpdf(s[0]) + st.norm(0, sig).logpdf(np.diff(s)).sum()
            + st.t(nu, 0, np.exp(s)).logpdf(r).sum())
    assert_allclose(u, -want, rtol=1e-12)


def test_hier_grad_fd():
    rng = np.random.RandomState(3)
    X = rng.normal(size=(40, 5))
    y = (rng.uniform(size=40) < 0.5).astype(float)
    z = rng.normal(size=7)
    u, g = m.hier_potential_grad(z, X, y)
    fd = m.central_fd(lambda zz: m.hier_potential(zz, X, y), z)
    assert_allclose(g, fd, rtol=1e-6, atol=1e-6)
    tau = np.exp(z[6])
    beta = z[5] + tau * z[:5]
    want = (st.norm().logpdf(z[:5]).sum() + st.norm().logpdf(z[5])
            + st.halfcauchy().logpdf(tau) + z[6]
            + st.bernoulli(1 / (1 + np.exp(-X @ beta))).logpmf(y).sum())
    assert_allclose(u, -want, rtol=1e-12)


def test_sv_noncentred_grad_fd_and_equivalence():
    rng = np.random.RandomState(4)
    T = 15
    r = rng.standard_t(4, size=T) * 0.02
    zeta = rng.normal(size=T)
    u_nu, u_sig = np.log(6.0), np.log(0.07)
    z = np.concatenate([[u_nu, u_sig], zeta])
    u, g = m.sv_nc_potential_grad(z, r)
    fd = m.central_fd(lambda zz: m.sv_nc_potential(zz, r), z)
    assert_allclose(g, fd, rtol=1e-5, atol=1e-5)
    # same density as the centred model up to the Jacobian of s = sigma*cumsum(zeta)
    s = np.exp(u_sig) * np.cumsum(zeta)
    zc = np.concatenate([[u_nu], s, [u_sig]])
    log_jac = T * u_sig  # |ds/dzeta| = sigma^T
    assert_allclose(u, m.sv_potential(zc, r) - log_jac, rtol=1e-12)


def test_digamma_reference_values():
    from scipy.special import digamma
    # psi(1) = -gamma, psi(1/2) = -gamma - 2 ln 2 (closed forms)
    assert_allclose(digamma(1.0), -0.5772156649015329, rtol=1e-14)
    assert_allclose(digamma(0.5), -0.5772156649015329 - 2 * np.log(2), rtol=1e-14)


def test_softplus_potentials_are_the_exp_potentials_under_the_change_of_variables():
    # With x = softplus(u) = exp(v), v = log softplus(u): the two unconstrained
    # potentials differ exactly by the two log-Jacobians,
    # U_softplus(u) = U_exp(v) + v + softplus(-u)   (transforms.py:606-652, 1156-1175)
    rng = np.random.default_rng(3)
    T = 40
    r = rng.normal(size=T) * 0.02
    for fn, pos in ((m.sv_potential, [0, T + 1]), (m.sv_nc_potential, [0, 1])):
        for _ in range(3):
            z = rng.normal(size=T + 2) * 0.5
            z[pos] = rng.normal(size=2)
            v = z.copy()
            v[pos] = np.log(d.softplus(z[pos]))
            want = fn(v, r) + np.sum(v[pos]) + np.sum(d.softplus(-z[pos]))
            assert_allclose(fn(z, r, positive="softplus"), want, rtol=1e-12)
    X = rng.normal(size=(30, 4))
    y = (rng.uniform(size=30) < 0.5).astype(np.float64)
    z = rng.normal(size=6)
    v = z.copy()
    v[5] = np.log(d.softplus(z[5]))
    assert_allclose(m.hier_potential(z, X, y, positive="softplus"),
                    m.hier_potential(v, X, y) + v[5] + d.softplus(-z[5]), rtol=1e-12)
