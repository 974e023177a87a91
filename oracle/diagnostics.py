"""ESS and (split) R-hat restated in numpy (oracle).

Restates ``numpyro/diagnostics.py:29-203``.  Pinned by the reference KATs
``test/test_diagnostics.py:48-108`` re-expressed in tests/test_oracle_diag.py.
"""
import numpy as np


def _compute_chain_variance_stats(x):
    """``diagnostics.py:29-42``."""
    C, N = x.shape[:2]
    chain_var = x.var(axis=1, ddof=1)
    var_within = chain_var.mean(axis=0)
    var_estimator = var_within * (N - 1) / N
    if x.shape[0] > 1:
        chain_mean = x.mean(axis=1)
        var_between = chain_mean.var(axis=0, ddof=1)
        var_estimator = var_estimator + var_between
    else:
        var_within = var_estimator
    return var_within, var_estimator


def gelman_rubin(x):
    """``diagnostics.py:45-61``."""
    var_within, var_estimator = _compute_chain_variance_stats(x)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.sqrt(var_estimator / var_within)


def split_gelman_rubin(x):
    """``diagnostics.py:64-80``."""
    n_half = x.shape[1] // 2
    return gelman_rubin(np.concatenate([x[:, :n_half], x[:, -n_half:]], axis=0))


def fft_next_fast_len(target):
    """``diagnostics.py:83-98``."""
    if target <= 2:
        return target
    while True:
        m = target
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            return target
        target += 1


def autocorrelation(x, axis=0, bias=True):
    """``diagnostics.py:101-142``."""
    N = x.shape[axis]
    M2 = 2 * fft_next_fast_len(N)
    x = np.swapaxes(x, axis, -1)
    c = x - x.mean(axis=-1, keepdims=True)
    f = np.fft.rfft(c, n=M2, axis=-1)
    ac = np.fft.irfft(f * np.conjugate(f), n=M2, axis=-1)[..., :N]
    if not bias:
        ac = ac / np.arange(N, 0.0, -1)
    with np.errstate(invalid="ignore", divide="ignore"):
        ac = (ac / ac[..., :1]).astype(np.float64)
    return np.swapaxes(ac, axis, -1)


def autocovariance(x, axis=0, bias=True):
    """``diagnostics.py:145-155``."""
    return autocorrelation(x, axis, bias) * x.var(axis=axis, keepdims=True)


def effective_sample_size(x, bias=True):
    """``diagnostics.py:158-203`` (Geyer initial monotone sequence)."""
    x = np.asarray(x)
    gamma_k_c = autocovariance(x, axis=1, bias=bias)
    var_within, var_estimator = _compute_chain_variance_stats(x)
    rho_k = 1.0 - (var_within - gamma_k_c.mean(axis=0)) / var_estimator
    rho_k[0] = 1.0
    Rho_k = rho_k[:-1:2, ...] + rho_k[1::2, ...]
    Rho_k = np.concatenate(
        [Rho_k[:1], np.minimum.accumulate(np.clip(Rho_k[1:, ...], 0, None), axis=0)], axis=0)
    tau = -1.0 + 2.0 * Rho_k.sum(axis=0)
    return np.prod(x.shape[:2]) / tau
