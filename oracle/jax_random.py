"""jax.random semantics used by the sampler, restated in numpy (oracle).

jax / jaxlib (>=0.7.0, ``pyproject.toml:22-24`` of the reference) are absent
here, so this module restates the published algorithm:

* threefry2x32 with 20 rounds (Salmon et al., "Parallel random numbers: as
  easy as 1, 2, 3", SC'11).  Pinned by the Random123 known-answer vectors
  (tests/test_oracle_random.py).
* key split / 32-bit random bits in the two key-derivation modes of jax
  ("partitionable", jax's default since 0.5, and "original").
* uniform (mantissa trick), normal (sqrt(2) * erfinv(u)), bernoulli (u < p).

Call sites in the reference: ``numpyro/infer/hmc.py:102`` (normal),
``hmc.py:335,472`` (split 3), ``numpyro/infer/hmc_util.py:565,656,920,1005,
1161`` (split), ``hmc_util.py:804,1162`` (bernoulli), ``numpyro/infer/mcmc.py
:671`` (split C).

Parity of the key-derivation mode against jax itself is unpinned here (jax
is absent); the threefry block cipher is pinned by the published KATs.
"""
import numpy as np
from scipy.special import erfinv

_ROT = ((13, 15, 26, 6), (17, 29, 16, 24))
_PARITY = np.uint32(0x1BD11BDA)

# The default derivation mode of jax>=0.5 (jax_threefry_partitionable=True).
PARTITIONABLE = True


def _rotl(x, r):
    return (x << np.uint32(r)) | (x >> np.uint32(32 - r))


def threefry2x32(key, x0, x1):
    """threefry2x32-20 block cipher, elementwise over uint32 arrays x0, x1."""
    k0 = np.uint32(key[0])
    k1 = np.uint32(key[1])
    ks = (k0, k1, k0 ^ k1 ^ _PARITY)
    x0 = np.asarray(x0, dtype=np.uint32).copy()
    x1 = np.asarray(x1, dtype=np.uint32).copy()
    with np.errstate(over="ignore"):
        x0 += ks[0]
        x1 += ks[1]
        for g in range(5):
            for r in _ROT[g % 2]:
                x0 += x1
                x1 = _rotl(x1, r)
                x1 ^= x0
            x0 += ks[(g + 1) % 3]
            x1 += ks[(g + 2) % 3] + np.uint32(g + 1)
    return x0, x1


def _tf_count(key, count):
    """jax's ``threefry_2x32(key, count)``: halves of the (padded) count."""
    count = np.asarray(count, dtype=np.uint32).ravel()
    n = count.size
    if n % 2:
        count = np.concatenate([count, np.zeros(1, np.uint32)])
    h = count.size // 2
    o0, o1 = threefry2x32(key, count[:h], count[h:])
    return np.concatenate([o0, o1])[:n]


def PRNGKey(seed):
    """``random.key(seed)`` / ``PRNGKey(seed)`` for 0 <= seed < 2**32."""
    return np.array([0, seed & 0xFFFFFFFF], dtype=np.uint32)


def split(key, num=2, partitionable=None):
    """``random.split(key, num)`` -> uint32 array [num, 2].  Partitionable
    mode: parity unpinned against jax (see the module docstring)."""
    part = PARTITIONABLE if partitionable is None else partitionable
    key = np.asarray(key, dtype=np.uint32)
    if part:
        o0, o1 = threefry2x32(key, np.zeros(num, np.uint32), np.arange(num, dtype=np.uint32))
        return np.stack([o0, o1], axis=-1)
    out = _tf_count(key, np.arange(2 * num, dtype=np.uint32))
    return out.reshape(num, 2)


def random_bits(key, n, partitionable=None):
    """32-bit ``random_bits(key, 32, shape)`` for a flat shape of ``n``
    elements; ``n=None`` means the scalar shape ``()``."""
    part = PARTITIONABLE if partitionable is None else partitionable
    key = np.asarray(key, dtype=np.uint32)
    if part:
        if n is None:
            o0, o1 = threefry2x32(key, np.zeros(1, np.uint32), np.zeros(1, np.uint32))
            return (o0 ^ o1)[0]
        o0, o1 = threefry2x32(key, np.zeros(n, np.uint32), np.arange(n, dtype=np.uint32))
        return o0 ^ o1
    if n is None:
        return _tf_count(key, np.zeros(1, np.uint32))[0]
    return _tf_count(key, np.arange(n, dtype=np.uint32))


def bits_to_unit_float32(bits):
    """float32 in [0, 1): (bits >> 9) | 0x3F800000 reinterpreted, minus 1."""
    b = (np.asarray(bits, dtype=np.uint32) >> np.uint32(9)) | np.uint32(0x3F800000)
    return b.view(np.float32) - np.float32(1.0)


def uniform(key, n=None, minval=0.0, maxval=1.0, partitionable=None):
    """float32 ``random.uniform``: f * (max - min) + min, clamped below by min."""
    f = bits_to_unit_float32(random_bits(key, n, partitionable))
    lo = np.float32(minval)
    hi = np.float32(maxval)
    out = f * (hi - lo) + lo
    return np.maximum(lo, out)


_NORMAL_LO = np.nextafter(np.float32(-1.0), np.float32(0.0))


def normal(key, n, partitionable=None):
    """``random.normal(key, (n,))`` = sqrt(2) * erfinv(u), u ~ U(nextafter(-1,0), 1)
    in float32; erfinv evaluated in float64 (the mathematical value).  The
    float32 bits of XLA's erf_inv: parity unpinned."""
    u = uniform(key, n, _NORMAL_LO, 1.0, partitionable)
    return np.sqrt(2.0) * erfinv(u.astype(np.float64))


def bernoulli(key, p=0.5, partitionable=None):
    """scalar ``random.bernoulli(key, p)`` = uniform(key, ()) < p."""
    u = uniform(key, None, 0.0, 1.0, partitionable)
    return bool(np.float64(u) < np.float64(p))
