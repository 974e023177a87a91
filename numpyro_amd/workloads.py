"""Synthetic logistic-regression workloads.

Shapes follow the benchmark configurations:
  c1: X in R^{1000 x 25} N(0,1) + intercept, beta* ~ N(0,1); 4 chains.
  c2: N=581,012, D=54 (10 columns N(0,1), 44 binary Bernoulli(0.1)) +
      intercept, beta* ~ N(0, 0.5^2); 256 chains.
  c5: N=2,000,000, D=256, X ~ N(0,1) with 51 (20%) columns scaled x10,
      beta* ~ N(0, 0.1^2); 1024 chains.
y ~ Bernoulli(sigmoid(X beta*)).  The draws come from numpy's PCG64
(``default_rng(seed)``, one seed per configuration), in float64 and in the
order listed, X then cast to float32; the logits are computed in float64
from the float32 X.  ``design`` returns the numpy arrays, so a CPU checker
regenerates a design bit for bit; ``make`` moves them to the GPU.
"""
import numpy as np
import torch

CONFIGS = {
    "c1": dict(N=1000, D=26, chains=4, seed=1),
    "c2": dict(N=581_012, D=55, chains=256, seed=2),
    "c5": dict(N=2_000_000, D=256, chains=1024, seed=5),
}

_ROWS = 1 << 17  # rows per generation chunk (float64 scratch of 268 MB at D=256)


def _normal_rows(rng, N, D):
    X = np.empty((N, D), np.float32)
    for r0 in range(0, N, _ROWS):
        X[r0:r0 + _ROWS] = rng.standard_normal((min(_ROWS, N - r0), D))
    return X


def _labels(rng, X, beta):
    N = X.shape[0]
    logits = np.empty(N)
    for r0 in range(0, N, _ROWS):
        logits[r0:r0 + _ROWS] = X[r0:r0 + _ROWS].astype(np.float64) @ beta
    return (rng.uniform(size=N) < 1.0 / (1.0 + np.exp(-logits))).astype(np.float32)


def design(name, N=None):
    """(X float32 [N, D], y float32 [N], beta* float64 [D], cfg) as numpy
    arrays, drawn from default_rng(cfg['seed'])."""
    cfg = dict(CONFIGS[name])
    if N is not None:
        cfg["N"] = int(N)
    N, D = cfg["N"], cfg["D"]
    rng = np.random.default_rng(cfg["seed"])
    if name == "c1":
        X = np.concatenate([_normal_rows(rng, N, D - 1), np.ones((N, 1), np.float32)], 1)
        beta = rng.standard_normal(D)
    elif name == "c2":
        Xc = _normal_rows(rng, N, 10)
        Xb = (rng.uniform(size=(N, 44)) < 0.1).astype(np.float32)
        X = np.concatenate([Xc, Xb, np.ones((N, 1), np.float32)], 1)
        beta = 0.5 * rng.standard_normal(D)
    elif name == "c5":
        X = _normal_rows(rng, N, D)
        cols = rng.choice(D, 51, replace=False)
        X[:, cols] *= 10.0
        beta = 0.1 * rng.standard_normal(D)
    else:
        raise ValueError(name)
    y = _labels(rng, X, beta)
    return X, y, beta, cfg


def make(name, device="cuda", N=None, chains=None):
    """``design`` moved to `device` as torch tensors (beta* float32)."""
    X, y, beta, cfg = design(name, N)
    if chains is not None:
        cfg["chains"] = int(chains)
    return (torch.from_numpy(X).to(device), torch.from_numpy(y).to(device),
            torch.from_numpy(beta.astype(np.float32)).to(device), cfg)


def make_sv(T=3000, seed=3):
    """Returns generated from the stochastic-volatility model itself:
    sigma* ~ Exponential(50), nu* ~ Exponential(0.1), s_0 ~ N(0, sigma*),
    s_t = s_{t-1} + sigma* z_t, r_t = exp(s_t) * StudentT(nu*).  numpy PCG64."""
    rng = np.random.default_rng(seed)
    sigma = rng.exponential(scale=1 / 50.0)
    nu = rng.exponential(scale=1 / 0.1)
    s = np.cumsum(sigma * rng.normal(size=T))
    r = np.exp(s) * rng.standard_t(max(nu, 1e-3), size=T)
    return r.astype(np.float32), dict(sigma=sigma, nu=nu, s=s, T=T, seed=seed)


def make_hier(G=1024, seed=4, nmin=50, nmax=5000, dmin=4, dmax=64):
    """Ragged hierarchical logistic workload: for each of G groups
    N_i ~ U{nmin..nmax}, D_i ~ U{dmin..dmax}, X_i ~ N(0,1), mu_i ~ N(0,1),
    tau_i ~ HalfCauchy(1), eta_i ~ N(0,1)^D_i, beta_i = mu_i + tau_i eta_i,
    y_i ~ Bernoulli(sigmoid(X_i beta_i)).  numpy PCG64."""
    rng = np.random.default_rng(seed)
    Xs, ys = [], []
    for _ in range(G):
        n = int(rng.integers(nmin, nmax + 1))
        d = int(rng.integers(dmin, dmax + 1))
        X = rng.normal(size=(n, d)).astype(np.float32)
        mu = rng.normal()
        tau = abs(rng.standard_cauchy())
        beta = mu + tau * rng.normal(size=d)
        y = (rng.uniform(size=n) < 1 / (1 + np.exp(-np.clip(X @ beta, -60, 60)))).astype(np.float32)
        Xs.append(X)
        ys.append(y)
    return Xs, ys


def laplace(X, y, prior_scale=1.0, iters=12, chunk=1 << 18, full_cov=False):
    """Mode and covariance of the Laplace approximation of the logistic
    posterior: Newton iterations on U(b) = -sum log Bernoulli(y |
    sigmoid(X b)) + |b|^2 / (2 s^2), the Hessian X^T diag(w) X + I / s^2
    accumulated in float64 over row chunks.  Returns (mode, diag(H^-1)), or
    (mode, H^-1) with ``full_cov``, as float64 tensors on X's device."""
    N, D = X.shape
    b = torch.zeros(D, dtype=torch.float64, device=X.device)
    eye = torch.eye(D, dtype=torch.float64, device=X.device) / prior_scale ** 2
    for _ in range(iters):
        g = b / prior_scale ** 2
        H = eye.clone()
        for r0 in range(0, N, chunk):
            xc = X[r0:r0 + chunk].to(torch.float64)
            p = torch.sigmoid(xc @ b)
            g += xc.T @ (p - y[r0:r0 + chunk].to(torch.float64))
            H += xc.T @ (xc * (p * (1 - p))[:, None])
        step = torch.linalg.solve(H, g)
        b = b - step
        if float(step.abs().max()) < 1e-10:
            break
    cov = torch.linalg.inv(H)
    return b, (0.5 * (cov + cov.T) if full_cov else torch.diagonal(cov).clone())


def laplace_init(X, y, C, seed=0, full_cov=False, return_mode=False):
    """init_params (mode + sd * uniform(-2, 2) per coordinate and chain) and
    the initial inverse_mass_matrix (Laplace variances, or the Laplace
    covariance with ``full_cov``), both float32; with ``return_mode`` also
    the float64 mode and standard deviations."""
    mode, cov = laplace(X, y, full_cov=full_cov)
    var = torch.diagonal(cov) if full_cov else cov
    sd = var.sqrt()
    u = torch.as_tensor(np.random.default_rng(seed).uniform(-2.0, 2.0, (C, X.shape[1])),
                        device=X.device)
    z0 = (mode[None, :] + sd[None, :] * u).to(torch.float32).contiguous()
    if return_mode:
        return z0, cov.to(torch.float32).contiguous(), mode, sd
    return z0, cov.to(torch.float32).contiguous()
