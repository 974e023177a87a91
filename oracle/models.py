"""Potentials U(z) = -log p(z) (unconstrained) and analytic gradients, float64.

U follows ``numpyro/infer/util.py:300-358`` (``_unconstrain_reparam`` +
``potential_energy``): real-support sites pass through, positive-support
sites get ``biject_to(positive)`` = ``ExpTransform`` (``numpyro/distributions/
transforms.py:2115-2118``) and add log|J| = z.  Latents are flattened in
sorted site-name order (``ravel_pytree`` of a dict).

Pins: U against a direct sum of ``oracle.distributions`` log-probs (which
are pinned to scipy) and the gradients against central finite differences
(as ``test/test_distributions.py:2155-2229``).
"""
import numpy as np
from scipy.special import digamma, expit

from . import distributions as dist


# ---- logistic regression (examples/covtype.py:66-71, test_mcmc.py:104-168) --

def logistic_potential(beta, X, y, prior_scale=1.0):
    """U(beta) = -sum_n log Bern(y_n | logits=x_n.beta) - sum_d log N(beta_d|0,s)."""
    X = np.asarray(X, dtype=np.float64)
    beta = np.asarray(beta, dtype=np.float64)
    logits = X @ beta
    ll = np.sum(dist.bernoulli_logits_log_prob(logits, y))
    lp = np.sum(dist.normal_log_prob(beta, 0.0, prior_scale))
    return -(ll + lp)


def logistic_potential_grad(beta, X, y, prior_scale=1.0):
    X = np.asarray(X, dtype=np.float64)
    beta = np.asarray(beta, dtype=np.float64)
    logits = X @ beta
    u = logistic_potential(beta, X, y, prior_scale)
    g = X.T @ (expit(logits) - y) + beta / (prior_scale * prior_scale)
    return u, g


# ---- stochastic volatility (examples/stochastic_volatility.py:58-66) ---------
# z = [log nu, s_0..s_{T-1}, log sigma]  (sorted: "nu", "s", "sigma")

SV_SIGMA_RATE = 50.0
SV_NU_RATE = 0.1


def _positive(u, positive):
    """(x, log|dx/du|) of a positive site: ExpTransform (transforms.py:606-652)
    or SoftplusTransform (transforms.py:1156-1175)."""
    if positive == "exp":
        return np.exp(u), u
    if positive == "softplus":
        return dist.softplus_transform(u)
    raise ValueError(positive)


def sv_potential(z, returns, sigma_rate=SV_SIGMA_RATE, nu_rate=SV_NU_RATE, positive="exp"):
    z = np.asarray(z, dtype=np.float64)
    T = returns.shape[0]
    u_nu, s, u_sig = z[0], z[1:1 + T], z[1 + T]
    (nu, j_nu), (sig, j_sig) = _positive(u_nu, positive), _positive(u_sig, positive)
    lp = dist.exponential_log_prob(sig, sigma_rate) + j_sig
    lp += dist.gaussian_random_walk_log_prob(s, sig)
    lp += dist.exponential_log_prob(nu, nu_rate) + j_nu
    lp += np.sum(dist.student_t_log_prob(returns, nu, 0.0, np.exp(s)))
    return -lp


def sv_potential_grad(z, returns, sigma_rate=SV_SIGMA_RATE, nu_rate=SV_NU_RATE):
    z = np.asarray(z, dtype=np.float64)
    T = returns.shape[0]
    u_nu, s, u_sig = z[0], z[1:1 + T], z[1 + T]
    nu, sig = np.exp(u_nu), np.exp(u_sig)
    # GRW
    d = np.diff(s)
    S2 = s[0] ** 2 + np.sum(d * d)
    gs = np.zeros(T)
    gs[0] -= s[0] / sig ** 2
    gs[1:] -= d / sig ** 2
    gs[:-1] += d / sig ** 2
    g_usig = S2 / sig ** 2 - T - sigma_rate * sig + 1.0
    # StudentT likelihood with scale exp(s)
    y = returns * np.exp(-s)
    y2 = y * y
    gs += (nu + 1.0) * y2 / (nu + y2) - 1.0
    dnu = np.sum(-0.5 * np.log1p(y2 / nu) + 0.5 * (nu + 1.0) * y2 / (nu * (nu + y2))
                 - 0.5 / nu - 0.5 * digamma(0.5 * nu) + 0.5 * digamma(0.5 * (nu + 1.0)))
    g_unu = nu * dnu - nu_rate * nu + 1.0
    g = np.concatenate([[g_unu], gs, [g_usig]])
    return sv_potential(z, returns, sigma_rate, nu_rate), -g


# ---- non-centred hierarchical logistic regression, one group ----------------
# z = [eta_0..eta_{D-1}, mu, log tau]  (sorted: "eta", "mu", "tau")

def hier_potential(z, X, y, positive="exp"):
    z = np.asarray(z, dtype=np.float64)
    D = X.shape[1]
    eta, mu, u_tau = z[:D], z[D], z[D + 1]
    tau, j_tau = _positive(u_tau, positive)
    beta = mu + tau * eta
    lp = np.sum(dist.normal_log_prob(eta)) + dist.normal_log_prob(mu)
    lp += dist.half_cauchy_log_prob(tau, 1.0) + j_tau
    lp += np.sum(dist.bernoulli_logits_log_prob(np.asarray(X, np.float64) @ beta, y))
    return -lp


def hier_potential_grad(z, X, y):
    z = np.asarray(z, dtype=np.float64)
    X = np.asarray(X, dtype=np.float64)
    D = X.shape[1]
    eta, mu, u_tau = z[:D], z[D], z[D + 1]
    tau = np.exp(u_tau)
    beta = mu + tau * eta
    gb = X.T @ (y - expit(X @ beta))          # d loglik / d beta
    g_eta = -eta + tau * gb
    g_mu = -mu + np.sum(gb)
    g_utau = tau * np.dot(eta, gb) - 2.0 * tau * tau / (1.0 + tau * tau) + 1.0
    g = np.concatenate([g_eta, [g_mu, g_utau]])
    return hier_potential(z, X, y), -g


# ---- finite differences (test helper) --------------------------------------

def central_fd(f, z, eps=1e-5):
    z = np.asarray(z, dtype=np.float64)
    g = np.zeros_like(z)
    for i in range(z.size):
        e = np.zeros_like(z)
        e[i] = eps
        g[i] = (f(z + e) - f(z - e)) / (2 * eps)
    return g


# ---- stochastic volatility, non-centred: s = sigma * cumsum(zeta) ----------
# z = [log nu, log sigma, zeta_0..zeta_{T-1}]  (sorted: "nu", "sigma", "zeta");
# zeta ~ N(0, 1) i.i.d. gives the same law of s as the Gaussian random walk
# of examples/stochastic_volatility.py:58-66 (s_0 ~ N(0, sigma), increments
# N(0, sigma)).

def sv_nc_potential(z, returns, sigma_rate=SV_SIGMA_RATE, nu_rate=SV_NU_RATE, positive="exp"):
    z = np.asarray(z, dtype=np.float64)
    T = returns.shape[0]
    u_nu, u_sig, zeta = z[0], z[1], z[2:2 + T]
    (nu, j_nu), (sig, j_sig) = _positive(u_nu, positive), _positive(u_sig, positive)
    s = sig * np.cumsum(zeta)
    lp = dist.exponential_log_prob(sig, sigma_rate) + j_sig
    lp += np.sum(dist.normal_log_prob(zeta))
    lp += dist.exponential_log_prob(nu, nu_rate) + j_nu
    lp += np.sum(dist.student_t_log_prob(returns, nu, 0.0, np.exp(s)))
    return -lp


def sv_nc_potential_grad(z, returns, sigma_rate=SV_SIGMA_RATE, nu_rate=SV_NU_RATE):
    z = np.asarray(z, dtype=np.float64)
    T = returns.shape[0]
    u_nu, u_sig, zeta = z[0], z[1], z[2:2 + T]
    nu, sig = np.exp(u_nu), np.exp(u_sig)
    csum = np.cumsum(zeta)
    s = sig * csum
    y2 = (returns * np.exp(-s)) ** 2
    a = 1.0 - (nu + 1.0) * y2 / (nu + y2)           # dU/ds_t
    suffix = np.cumsum(a[::-1])[::-1]                 # sum_{tau >= t} a_tau
    g_zeta = zeta + sig * suffix
    g_usig = sig * np.sum(a * csum) + sigma_rate * sig - 1.0
    dnu = np.sum(0.5 * np.log1p(y2 / nu) - 0.5 * (nu + 1.0) * y2 / (nu * (nu + y2))
                 + 0.5 / nu + 0.5 * digamma(0.5 * nu) - 0.5 * digamma(0.5 * (nu + 1.0)))
    g_unu = nu * dnu + nu_rate * nu - 1.0
    return sv_nc_potential(z, returns, sigma_rate, nu_rate), np.concatenate([[g_unu, g_usig], g_zeta])
