"""``MCMCKernel`` and the ``MCMC`` driver (``numpyro/infer/mcmc.py``).

All chains of one run live in one engine handle on the GPU (the reference's
``chain_method="vectorized"``); "parallel" and "sequential" are accepted and
run the same way, since one GPU holds every chain.  Warmup and sampling run
on the device in one call with device-side collection (``fori_collect``
semantics, ``numpyro/util.py:321-454``).
"""
from abc import ABC, abstractmethod
import warnings

import numpy as np
import torch

from .. import random as nrandom
from ..diagnostics import print_summary


class MCMCKernel(ABC):
    """``numpyro/infer/mcmc.py:33-159``."""

    def postprocess_fn(self, model_args, model_kwargs):
        return lambda x: x

    @abstractmethod
    def init(self, rng_key, num_warmup, init_params, model_args, model_kwargs):
        raise NotImplementedError

    @abstractmethod
    def sample(self, state, model_args, model_kwargs):
        raise NotImplementedError

    @property
    def sample_field(self):
        raise NotImplementedError

    @property
    def default_fields(self):
        return (self.sample_field,)

    @property
    def is_ensemble_kernel(self):
        return False

    def get_diagnostics_str(self, state):
        return ""


class MCMC:
    """``numpyro/infer/mcmc.py:225-809``."""

    def __init__(self, sampler, *, num_warmup, num_samples, num_chains=1, thinning=1,
                 postprocess_fn=None, chain_method="parallel", progress_bar=True,
                 progress_rate=None, jit_model_args=False):
        self.sampler = sampler
        self._sample_field = sampler.sample_field
        self._default_fields = sampler.default_fields
        self.num_warmup = num_warmup
        self.num_samples = num_samples
        self.num_chains = num_chains
        if not isinstance(thinning, int) or thinning < 1:
            raise ValueError("thinning must be a positive integer")
        self.thinning = thinning
        self.postprocess_fn = postprocess_fn
        if not callable(chain_method) and chain_method not in [
                "parallel", "vectorized", "sequential"]:
            raise ValueError(
                "Only supporting the following methods to draw chains:"
                ' "sequential", "parallel", or "vectorized"')
        self.chain_method = chain_method
        self.progress_bar = progress_bar
        self._states = None
        self._states_flat = None
        self._last_state = None
        self._warmup_state = None
        self._grad_evals = 0
        self._collection_params = {}
        self._set_collection_params()

    def _set_collection_params(self, lower=None, upper=None, collection_size=None, phase=None):
        self._collection_params["lower"] = self.num_warmup if lower is None else lower
        self._collection_params["upper"] = (self.num_warmup + self.num_samples
                                            if upper is None else upper)
        self._collection_params["collection_size"] = collection_size
        self._collection_params["phase"] = phase

    @property
    def post_warmup_state(self):
        return self._warmup_state

    @post_warmup_state.setter
    def post_warmup_state(self, state):
        self._warmup_state = state

    @property
    def last_state(self):
        return self._last_state

    @property
    def grad_evals(self):
        """Useful gradient evaluations (leapfrog steps) of the last run."""
        return self._grad_evals

    def warmup(self, rng_key, *args, extra_fields=(), collect_warmup=False, init_params=None,
               **kwargs):
        """``mcmc.py:589-633``."""
        self._warmup_state = None
        if collect_warmup:
            self._set_collection_params(0, self.num_warmup, self.num_warmup, "warmup")
        else:
            self._set_collection_params(self.num_warmup, self.num_warmup, self.num_samples,
                                        "warmup")
        self.run(rng_key, *args, extra_fields=extra_fields, init_params=init_params, **kwargs)
        self._warmup_state = self._last_state

    def run(self, rng_key, *args, extra_fields=(), init_params=None, **kwargs):
        """``mcmc.py:635-729``."""
        collect_fields = {}
        remove_sites = {}
        for f in (self._sample_field,) + tuple(self._default_fields) + tuple(extra_fields):
            if f.startswith(f"~{self._sample_field}."):
                remove_sites[f[len(self._sample_field) + 2:]] = None
            else:
                collect_fields[f] = None
        collect_fields = tuple(collect_fields)
        if init_params is not None and self.num_chains > 1 and not isinstance(init_params, dict):
            if torch.as_tensor(init_params).shape[0] != self.num_chains:
                raise ValueError("`init_params` must have the same leading dimension"
                                 " as `num_chains`.")
        rng_key = np.asarray(rng_key, dtype=np.uint32)
        if rng_key.ndim == 1 and self.num_chains > 1:
            chain_keys = nrandom.split(rng_key, self.num_chains)
        else:
            chain_keys = rng_key.reshape(-1, 2)
        if self._warmup_state is not None:
            self._set_collection_params(0, self.num_samples, self.num_samples, "sample")
            eng = self.sampler._engine
            eng.set_state(self._warmup_state, rng_key=chain_keys)
        else:
            if init_params is not None and self.num_chains == 1 and \
                    not isinstance(init_params, dict):
                init_params = torch.as_tensor(init_params).reshape(1, -1)
            self.sampler.init(chain_keys, self.num_warmup, init_params, args, kwargs)
            eng = self.sampler._engine
        lower = self._collection_params["lower"]
        upper = self._collection_params["upper"]
        size = self._collection_params["collection_size"]
        size = (upper - lower) // self.thinning if size is None else size
        i0 = int(eng.get_state().i.max())
        start_idx = i0 + lower + (upper - lower) % self.thinning
        # HMCECS keeps the inner kernel's state under hmc_state and its own
        # Gibbs acceptance probability as accept_prob (HMCECSState,
        # numpyro/infer/hmc_gibbs.py:490-491, 660, 681); no other field exists
        ren = {}
        if getattr(self.sampler, "_gibbs_sites", None):
            for f in collect_fields:
                if f.startswith("hmc_state."):
                    ren[f] = f[len("hmc_state."):]
                elif f == "accept_prob":
                    ren[f] = "_gibbs_accept_prob"
                elif f != self._sample_field:
                    raise AttributeError(f"'HMCECSState' object has no attribute {f!r}")
        dev_fields = tuple(ren.get(f, f) for f in collect_fields if f != "z") + ("z",)
        out = eng.run(i0 + upper, collect_fields=dev_fields, start_idx=start_idx,
                      thinning=self.thinning, size=size)
        self._grad_evals = eng.grad_evals
        pp = self.postprocess_fn or self.sampler.postprocess_fn(args, kwargs)
        states = {}
        for f in collect_fields:
            if f == self._sample_field:
                z = pp(out["z"])
                gsites = getattr(self.sampler, "_gibbs_sites", ())
                if "_gibbs" in out and isinstance(z, dict) and gsites:
                    # HMCECS: z = {**z_gibbs, **hmc_state.z} (hmc_gibbs.py:680)
                    z[gsites[0]] = out["_gibbs"]
                if isinstance(z, dict):
                    for s in remove_sites:
                        z.pop(s, None)
                states[f] = z
            elif ren.get(f, f) in out:
                states[f] = out[ren.get(f, f)]
            else:
                raise ValueError(f"field {f!r} cannot be collected")
        self._states = states
        self._states_flat = None
        self._last_state = eng.get_state()
        self._set_collection_params()

    def _flat(self, x):
        if isinstance(x, dict):
            return {k: self._flat(v) for k, v in x.items()}
        return x.reshape((x.shape[0] * x.shape[1],) + tuple(x.shape[2:]))

    def get_samples(self, group_by_chain=False):
        """``mcmc.py:731-755``."""
        s = self._states[self._sample_field]
        return s if group_by_chain else self._flat(s)

    def get_extra_fields(self, group_by_chain=False):
        """``mcmc.py:757-767``."""
        return {k: (v if group_by_chain else self._flat(v))
                for k, v in self._states.items() if k != self._sample_field}

    def transfer_states_to_host(self):
        """``mcmc.py:799-804``: moves the collected samples and extra fields
        from HBM to host memory (torch CPU tensors)."""
        def host(x):
            if isinstance(x, dict):
                return {k: host(v) for k, v in x.items()}
            return x.cpu() if torch.is_tensor(x) else x
        self._states = host(self._states)
        if self._states_flat is not None:
            self._states_flat = host(self._states_flat)

    def print_summary(self, prob=0.9, exclude_deterministic=True):
        """``mcmc.py:769-797``."""
        sites = self._states[self._sample_field]
        if torch.is_tensor(sites):
            sites = sites.cpu().numpy()
        elif isinstance(sites, dict):
            skip = set(getattr(self.sampler, "_deterministic_sites", ())) \
                if exclude_deterministic else set()
            skip |= set(getattr(self.sampler, "_gibbs_sites", ()))  # subsample indices
            sites = {k: v.cpu().numpy() for k, v in sites.items() if k not in skip}
        print_summary(sites, prob=prob)
        ex = self.get_extra_fields()
        if "diverging" in ex:
            print("Number of divergences: {}".format(int(ex["diverging"].sum())))

    def __getstate__(self):
        warnings.warn("MCMC state holds GPU handles; pickling drops the sampler engine")
        st = self.__dict__.copy()
        return st
