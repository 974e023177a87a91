"""Log-densities and bijections on the hot path, float64 numpy (oracle).

Each function restates the reference's formula; tests pin them against
scipy.stats (as ``test/test_distributions.py:467-529,1741-1803`` does) and
the transforms' log-det against log|dy/dx| by finite differences
(``test/test_transforms.py:285-385``).
"""
import numpy as np
from scipy.special import gammaln

LOG_2PI_HALF = 0.5 * np.log(2.0 * np.pi)


def normal_log_prob(x, loc=0.0, scale=1.0):
    """``numpyro/distributions/continuous.py:2975-2989``."""
    z = (x - loc) / scale
    return -0.5 * z * z - np.log(np.sqrt(2.0 * np.pi) * scale)


def student_t_log_prob(x, df, loc=0.0, scale=1.0):
    """``continuous.py:3229-3238``."""
    y = (x - loc) / scale
    z = (np.log(scale) + 0.5 * np.log(df) + 0.5 * np.log(np.pi)
         + gammaln(0.5 * df) - gammaln(0.5 * (df + 1.0)))
    return -0.5 * (df + 1.0) * np.log1p(y * y / df) - z


def cauchy_log_prob(x, loc=0.0, scale=1.0):
    """``continuous.py:390-406``."""
    return -np.log(np.pi) - np.log(scale) - np.log1p(((x - loc) / scale) ** 2)


def half_cauchy_log_prob(x, scale=1.0):
    """``continuous.py:1149-1150`` (Cauchy(0, scale) + log 2, x >= 0)."""
    lp = cauchy_log_prob(x, 0.0, scale) + np.log(2.0)
    return np.where(np.asarray(x) >= 0, lp, -np.inf)


def exponential_log_prob(x, rate):
    """``continuous.py:715-726``; support x >= 0."""
    lp = np.log(rate) - rate * x
    return np.where(np.asarray(x) >= 0, lp, -np.inf)


def gaussian_random_walk_log_prob(x, scale):
    """``continuous.py:1104-1112``: N(0,scale) at x[0] plus N(0,scale) of the
    increments x[t]-x[t-1]."""
    x = np.asarray(x, dtype=np.float64)
    init = normal_log_prob(x[0], 0.0, scale)
    steps = normal_log_prob(x[1:] - x[:-1], 0.0, scale)
    return init + np.sum(steps)


def softplus(x):
    return np.logaddexp(0.0, x)


def bernoulli_logits_log_prob(logits, y):
    """``numpyro/distributions/discrete.py:244-263`` ->
    ``numpyro/distributions/util.py:317-320``:
    -(max(l,0) + log1p(exp(-|l|)) - l*y)."""
    return -(np.clip(logits, 0, None) + np.log1p(np.exp(-np.abs(logits))) - logits * y)


# ---- bijections -------------------------------------------------------------

def exp_transform(x):
    """``numpyro/distributions/transforms.py:606-652``: y = exp(x),
    log|dy/dx| = x (``647-650``)."""
    return np.exp(x), x


def softplus_transform(x):
    """``transforms.py:1156-1175``: y = softplus(x), log|dy/dx| = -softplus(-x)."""
    return softplus(x), -softplus(-x)


def softplus_inv(y):
    """``transforms.py:1152-1153``."""
    return np.log(-np.expm1(-y)) + y
