"""HMCECS (energy-conserving subsampling) for logistic regression, float64.

Test infrastructure only: the product never imports this module.

Restates, for the covtype model (``examples/covtype.py:66-71`` with
``subsample_size``; ``examples/covtype.py:154-165`` runs it under HMCECS):

- ``randint``: ``jax.random.randint`` for 32-bit integers (jax 0.5
  ``jax/_src/random.py`` ``_randint``, a dependency absent from the
  reference tree): two 32-bit draws from ``split(key)``, combined as
  ``(hi % span) * (2**32 mod span) + lo % span`` in wrapping uint32, mod span;
- ``update_block``: ``numpyro/contrib/ecs_proxies.py:58-71`` (one block of
  ``block_size = (m - 1) // num_blocks + 1`` indices redrawn with
  replacement; the padded index vector is updated by a start-clamped
  ``dynamic_update_slice`` and cut back to ``m``);
- ``taylor_stats`` / ``proxy``: ``ecs_proxies.py:95-250`` for the logistic
  likelihood, whose per-row gradient and Hessian at the reference point are
  ``r_i x_i`` and ``-w_i x_i x_i^T`` with ``r_i = y_i - sigmoid(l_i)``,
  ``w_i = sigmoid(l_i) (1 - sigmoid(l_i))``; degree 1 drops the Hessian;
- ``loglik_estimate``: ``ecs_proxies.py:23-50`` (difference estimator
  ``proxy_all + n mean(d) - n^2 / (2 m) var(d)``) or, without a proxy, the
  plate's ``n / m`` scaling of the subsample's log-likelihood
  (``primitives.py:474-637``);
- ``potential_grad``: ``U = -log N(beta | 0, s) - loglik_estimate``;
- ``gibbs_step``: ``numpyro/infer/hmc_gibbs.py:640-681``.

Pins (tests/test_ecs_oracle.py): the Taylor terms against a second-order
expansion built from finite-difference derivatives of the per-row
log-likelihood (as ``test/infer/test_hmc_gibbs.py:333-399`` does with jax
derivatives); the estimator at the reference point equals the full
log-likelihood; unbiasedness of the mean term by enumerating every
subsample of size one; the gradient against central differences;
``randint`` against ``(hi * 2**32 + lo) mod span`` in exact integers for
spans below 2**16; the block layout against hand-derived cases.
Stream parity of ``randint`` with jax itself: parity unpinned (no jax here).
"""
import numpy as np
from scipy.special import expit

from . import distributions as dist
from .jax_random import random_bits, split, uniform


def randint(key, n, minval, maxval, partitionable=None):
    """``random.randint(key, shape, minval, maxval)`` for int32; ``n=None``
    is the scalar shape."""
    k1, k2 = split(key, 2, partitionable)
    hi = np.atleast_1d(random_bits(k1, n, partitionable)).astype(np.uint64)
    lo = np.atleast_1d(random_bits(k2, n, partitionable)).astype(np.uint64)
    span = np.uint64(max(int(maxval) - int(minval), 1) & 0xFFFFFFFF)
    mask = np.uint64(0xFFFFFFFF)
    mult = np.uint64(2 ** 16) % span
    mult = (mult * mult & mask) % span
    off = (((hi % span) * mult & mask) + lo % span) & mask
    off = off % span
    out = (np.int64(minval) + off.astype(np.int64)).astype(np.int32)
    return out[0] if n is None else out


def block_layout(m, num_blocks):
    """(block_size, pad) of ``ecs_proxies.py:61-62``."""
    bs = (m - 1) // num_blocks + 1
    pad = bs - (m - 1) % bs - 1
    return bs, pad


def update_block(key, idx, num_blocks, n, partitionable=None):
    """One block update of the subsample indices ``idx`` (int32 [m]) with
    ``key``; returns the new indices."""
    m = idx.shape[0]
    keys = split(key, 3, partitionable)
    subkey, block_key = keys[1], keys[2]
    bs, pad = block_layout(m, num_blocks)
    chosen = int(randint(block_key, None, 0, num_blocks, partitionable))
    new = randint(subkey, bs, 0, n, partitionable)
    padded = np.concatenate([idx.astype(np.int32), np.zeros(pad, np.int32)])
    start = min(max(chosen * bs, 0), padded.shape[0] - bs)
    padded[start:start + bs] = new
    return padded[:m]


def row_loglik(logits, y):
    return dist.bernoulli_logits_log_prob(logits, y)


def taylor_stats(X, y, beta_ref, degree=2):
    """Per-row (l_ref, ll_ref, r, w) and the full-data sums (LL, G, H)."""
    X = np.asarray(X, np.float64)
    y = np.asarray(y, np.float64)
    l_ref = X @ np.asarray(beta_ref, np.float64)
    p = expit(l_ref)
    ll_ref = row_loglik(l_ref, y)
    r = y - p
    w = p * (1.0 - p) if degree == 2 else np.zeros_like(p)
    G = X.T @ r
    H = -(X * w[:, None]).T @ X
    return dict(beta_ref=np.asarray(beta_ref, np.float64), l_ref=l_ref, ll_ref=ll_ref, r=r, w=w,
                LL=ll_ref.sum(), G=G, H=H)


def proxy(beta, idx, X, st):
    """(proxy over all rows, proxy of each subsample row) at ``beta``."""
    d = np.asarray(beta, np.float64) - st["beta_ref"]
    dl = np.asarray(X, np.float64)[idx] @ d
    sub = st["ll_ref"][idx] + st["r"][idx] * dl - 0.5 * st["w"][idx] * dl * dl
    full = st["LL"] + st["G"] @ d + 0.5 * d @ (st["H"] @ d)
    return full, sub


def loglik_estimate(beta, idx, X, y, n, st=None):
    X = np.asarray(X, np.float64)
    ll = row_loglik(X[idx] @ np.asarray(beta, np.float64), np.asarray(y, np.float64)[idx])
    m = idx.shape[0]
    if st is None:
        return n / m * ll.sum()
    full, sub = proxy(beta, idx, X, st)
    diff = ll - sub
    return full + n * diff.mean() - 0.5 * n * n / m * diff.var()


def potential(beta, idx, X, y, n, st=None, prior_scale=1.0):
    lp = np.sum(dist.normal_log_prob(np.asarray(beta, np.float64), 0.0, prior_scale))
    return -(lp + loglik_estimate(beta, idx, X, y, n, st))


def potential_grad(beta, idx, X, y, n, st=None, prior_scale=1.0):
    """U and dU/dbeta.  With a proxy, d diff_j / d l_j = (y - sigmoid(l)) -
    r + w (l - l_ref) and the estimator's weight on diff_j is
    (n / m) (1 - (n / m) (diff_j - mean diff))."""
    X = np.asarray(X, np.float64)
    y = np.asarray(y, np.float64)
    beta = np.asarray(beta, np.float64)
    m = idx.shape[0]
    Xs, ys = X[idx], y[idx]
    lg = Xs @ beta
    u = potential(beta, idx, X, y, n, st, prior_scale)
    if st is None:
        c = n / m * (ys - expit(lg))
        g_ll = Xs.T @ c
    else:
        d = beta - st["beta_ref"]
        dl = lg - st["l_ref"][idx]
        diff = row_loglik(lg, ys) - (st["ll_ref"][idx] + st["r"][idx] * dl
                                     - 0.5 * st["w"][idx] * dl * dl)
        cj = (ys - expit(lg)) - st["r"][idx] + st["w"][idx] * dl
        a = n / m * (1.0 - n / m * (diff - diff.mean()))
        g_ll = st["G"] + st["H"] @ d + Xs.T @ (a * cj)
    return u, beta / (prior_scale * prior_scale) - g_ll


def init_keys(key, proxy_on, partitionable=None):
    """Per-chain key derivations of ``hmc_gibbs.py:577-638`` and ``123-152``:
    returns (initial-index key, HMCECS state key, inner-kernel key)."""
    rng_key, key_u = split(key, 2, partitionable)
    if proxy_on:
        rng_key = split(rng_key, 2, partitionable)[0]
    rng_key, key_z = split(rng_key, 2, partitionable)
    return split(key_u, 2, partitionable)[1], rng_key, key_z


def initial_indices(key_idx, m, n, partitionable=None):
    """Initial subsample: m indices drawn with replacement."""
    return randint(key_idx, m, 0, n, partitionable)


def gibbs_step(key, idx, beta, pe, X, y, n, num_blocks, st=None, prior_scale=1.0,
               partitionable=None):
    """One Gibbs update of the subsample (``hmc_gibbs.py:640-672``).  Returns
    (new state key, idx, pe, grad or None if rejected, accept_prob)."""
    rng_key = split(key, 2, partitionable)[0]
    idx_new = update_block(rng_key, idx, num_blocks, n, partitionable)
    pe_new, g_new = potential_grad(beta, idx_new, X, y, n, st, prior_scale)
    accept_prob = min(np.exp(pe - pe_new), 1.0)
    u = np.float64(uniform(rng_key, None, 0.0, 1.0, partitionable))
    if u < accept_prob:
        return rng_key, idx_new, pe_new, g_new, accept_prob
    return rng_key, idx, pe, None, accept_prob
