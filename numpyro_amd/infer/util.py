"""``fori_collect`` (``numpyro/util.py:321-454``) for the engine's kernels.

For a ``sample_kernel`` returned by ``numpyro_amd.infer.hmc`` the whole loop
runs on the device with device-side collection; ``transform`` is then applied
to the collected states (it sees an ``HMCState``-like object whose ``z``
holds [num_chains, collection_size, dim])."""
from collections import namedtuple

import torch

_Collected = namedtuple("_Collected", ["z", "num_steps", "accept_prob", "potential_energy",
                                       "diverging"])


def fori_collect(lower, upper, body_fun, init_val, transform=lambda x: x, progbar=True,
                 return_last_val=False, collection_size=None, thinning=1, **progbar_opts):
    assert lower <= upper
    assert thinning >= 1
    size = (upper - lower) // thinning if collection_size is None else collection_size
    eng = body_fun.engine() if getattr(body_fun, "_nax", False) else None
    if eng is None:
        raise ValueError("fori_collect needs a sample_kernel from numpyro_amd.infer.hmc")
    i0 = int(torch.as_tensor(init_val.i).max())
    if int(eng.get_state().i.max()) != i0:
        eng.set_state(init_val)
    start_idx = i0 + lower + (upper - lower) % thinning
    out = eng.run(i0 + upper, collect_fields=("z", "num_steps", "accept_prob",
                                              "potential_energy", "diverging"),
                  start_idx=start_idx, thinning=thinning, size=size)
    coll = _Collected(out["z"], out["num_steps"], out["accept_prob"],
                      out["potential_energy"], out["diverging"])
    res = transform(coll)
    if eng.C == 1:
        res = _squeeze_chain(res)
    if return_last_val:
        return res, eng.get_state()
    return res


def _squeeze_chain(x):
    if isinstance(x, dict):
        return {k: _squeeze_chain(v) for k, v in x.items()}
    if isinstance(x, tuple):
        return type(x)(*[_squeeze_chain(v) for v in x])
    if torch.is_tensor(x) and x.ndim >= 1 and x.shape[0] == 1:
        return x[0]
    return x
