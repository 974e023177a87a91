"""jax.random-compatible keys for the engine.

A key is a numpy uint32 array of shape (2,) (``random.key(seed)`` /
``PRNGKey(seed)`` is the pair (0, seed)).  Key derivations run on the GPU
through the engine's threefry2x32 kernels (``nax_random_split``,
``nax_chain_keys``).
"""
import numpy as np
import torch

from . import _lib

PARTITIONABLE = True  # jax_threefry_partitionable, the default since jax 0.5


def PRNGKey(seed):
    return np.array([0, int(seed) & 0xFFFFFFFF], dtype=np.uint32)


key = PRNGKey


def _dev_u32(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.uint32))
    return torch.from_numpy(a.view(np.int32).copy()).cuda()


def _host_u32(t):
    return t.cpu().numpy().view(np.uint32).copy()


def is_prng_key(k):
    k = np.asarray(k)
    return k.shape == (2,) and k.dtype == np.uint32


def split(k, num=2, partitionable=None):
    part = PARTITIONABLE if partitionable is None else partitionable
    lib = _lib.load()
    dk = _dev_u32(k)
    out = torch.empty(num * 2, dtype=torch.int32, device="cuda")
    _lib.check(lib.nax_random_split(_lib.ptr(dk), num, int(part), _lib.ptr(out)), "split")
    return _host_u32(out).reshape(num, 2)


def normal(k, n, partitionable=None):
    """``random.normal(key, (n,))`` in float32, as a cuda tensor."""
    part = PARTITIONABLE if partitionable is None else partitionable
    lib = _lib.load()
    dk = _dev_u32(k)
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    _lib.check(lib.nax_random_normal(_lib.ptr(dk), n, int(part), _lib.ptr(out)), "normal")
    return out


def chain_init_keys(k, num_chains, partitionable=None):
    """Keys that MCMC.run + HMC.init hand to init_kernel per chain."""
    part = PARTITIONABLE if partitionable is None else partitionable
    lib = _lib.load()
    dk = _dev_u32(k)
    out = torch.empty(num_chains * 2, dtype=torch.int32, device="cuda")
    _lib.check(lib.nax_chain_keys(_lib.ptr(dk), num_chains, int(part), _lib.ptr(out)),
               "chain_keys")
    return _host_u32(out).reshape(num_chains, 2)
