"""Posterior diagnostics: ESS, (split) R-hat, HPDI, summary.

Same formulas as ``numpyro/diagnostics.py``: ESS by autocovariance and
Geyer's initial monotone sequence (``diagnostics.py:101-203``), split
Gelman-Rubin (``64-80``).  Host NumPy for arrays; ``device_diagnostics``
computes ESS and split R-hat of draws already on the GPU with the library's
kernel (one workgroup per parameter).
"""
import numpy as np


def _chain_var_stats(x):
    C, N = x.shape[:2]
    chain_var = x.var(axis=1, ddof=1)
    var_within = chain_var.mean(axis=0)
    var_estimator = var_within * (N - 1) / N
    if C > 1:
        var_estimator = var_estimator + x.mean(axis=1).var(axis=0, ddof=1)
    else:
        var_within = var_estimator
    return var_within, var_estimator


def gelman_rubin(x):
    x = np.asarray(x, dtype=np.float64)
    vw, ve = _chain_var_stats(x)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.sqrt(ve / vw)


def split_gelman_rubin(x):
    x = np.asarray(x, dtype=np.float64)
    h = x.shape[1] // 2
    return gelman_rubin(np.concatenate([x[:, :h], x[:, -h:]], axis=0))


def _next_fast_len(n):
    if n <= 2:
        return n
    while True:
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            return n
        n += 1


def autocorrelation(x, axis=0, bias=True):
    x = np.asarray(x, dtype=np.float64)
    N = x.shape[axis]
    M2 = 2 * _next_fast_len(N)
    x = np.swapaxes(x, axis, -1)
    c = x - x.mean(axis=-1, keepdims=True)
    f = np.fft.rfft(c, n=M2, axis=-1)
    ac = np.fft.irfft(f * np.conjugate(f), n=M2, axis=-1)[..., :N]
    if not bias:
        ac = ac / np.arange(N, 0.0, -1)
    with np.errstate(invalid="ignore", divide="ignore"):
        ac = ac / ac[..., :1]
    return np.swapaxes(ac, axis, -1)


def autocovariance(x, axis=0, bias=True):
    x = np.asarray(x, dtype=np.float64)
    return autocorrelation(x, axis, bias) * x.var(axis=axis, keepdims=True)


def effective_sample_size(x, bias=True):
    x = np.asarray(x, dtype=np.float64)
    gamma = autocovariance(x, axis=1, bias=bias)
    vw, ve = _chain_var_stats(x)
    rho = 1.0 - (vw - gamma.mean(axis=0)) / ve
    rho[0] = 1.0
    R = rho[:-1:2, ...] + rho[1::2, ...]
    R = np.concatenate([R[:1], np.minimum.accumulate(np.clip(R[1:, ...], 0, None), axis=0)],
                       axis=0)
    tau = -1.0 + 2.0 * R.sum(axis=0)
    return np.prod(x.shape[:2]) / tau


def hpdi(x, prob=0.90, axis=0):
    x = np.asarray(x, dtype=np.float64)
    x = np.swapaxes(x, axis, 0)
    sorted_x = np.sort(x, axis=0)
    mass = x.shape[0]
    idx = int(np.floor(prob * mass))
    lows = sorted_x[:mass - idx]
    highs = sorted_x[idx:]
    i = np.argmin(highs - lows, axis=0)
    lo = np.take_along_axis(lows, np.expand_dims(i, 0), 0)[0]
    hi = np.take_along_axis(highs, np.expand_dims(i, 0), 0)[0]
    return np.swapaxes(np.stack([lo, hi]), 0, axis)


def summary(samples, prob=0.90, group_by_chain=True):
    if not isinstance(samples, dict):
        samples = {"Param:{}".format(i): v for i, v in enumerate(
            [samples] if np.ndim(samples) else [])} or {"Param:0": samples}
    out = {}
    for name, v in samples.items():
        v = np.asarray(v, dtype=np.float64)
        if not group_by_chain:
            v = v[None]
        flat = v.reshape((-1,) + v.shape[2:])
        lo, hi = hpdi(flat, prob)
        out[name] = {"mean": flat.mean(0), "std": flat.std(0), "median": np.median(flat, 0),
                     f"{50 * (1 - prob):.1f}%": lo, f"{50 * (1 + prob):.1f}%": hi,
                     "n_eff": effective_sample_size(v),
                     "r_hat": split_gelman_rubin(v) if v.shape[1] >= 4 else np.nan}
    return out


def print_summary(samples, prob=0.90, group_by_chain=True):
    s = summary(samples, prob, group_by_chain)
    for name, st in s.items():
        mean = np.atleast_1d(st["mean"])
        print(f"{name}: mean[:4]={mean[:4]} min n_eff={np.min(st['n_eff']):.1f} "
              f"max r_hat={np.nanmax(np.atleast_1d(st['r_hat'])):.3f}")


def device_diagnostics(x):
    """(ESS, split R-hat) per parameter of draws ``x`` [chains, draws, ...] on
    the GPU (float32 is read; results float64, on the GPU).  draws <= 4096."""
    import torch

    from . import _lib
    x = torch.as_tensor(x)
    if x.ndim < 2:
        raise ValueError("x must be [chains, draws, ...]")
    C, S = int(x.shape[0]), int(x.shape[1])
    rest = tuple(x.shape[2:])
    P = 1
    for n in rest:
        P *= int(n)
    xf = x.to(device="cuda", dtype=torch.float32).reshape(C, S, P).contiguous()
    ess = torch.empty(P, dtype=torch.float64, device="cuda")
    rhat = torch.empty(P, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _lib.check(_lib.load().nax_diagnostics(_lib.ptr(xf), C, S, P, _lib.ptr(ess), _lib.ptr(rhat)),
               "nax_diagnostics")
    return ess.reshape(rest), rhat.reshape(rest)
