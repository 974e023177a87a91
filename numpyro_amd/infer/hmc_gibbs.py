"""HMCECS: HMC/NUTS within Gibbs over subsample indices (energy-conserving
subsampling), ``numpyro/infer/hmc_gibbs.py:502-689`` with the Taylor proxy
of ``numpyro/contrib/ecs_proxies.py:95-250``.

The subsample of every chain lives on the GPU ([C, m] int32).  Before each
iteration ``nax_ecs_gibbs`` redraws one block of each chain's indices,
evaluates U and dU at the chain's position with the proposal and accepts it
with probability min(1, exp(U - U_new)); the inner NUTS/HMC iteration then
runs on the chain's (possibly new) subsample.  Models: logistic regression
with a subsample plate over the rows of X (``examples/covtype.py:66-71`` with
``subsample_size``), traced by the model compiler.
"""
import numpy as np
import torch

from .. import _lib
from .. import random as nrandom
from ..compile import compile_model, is_model_function
from ..models import SubsampledLogisticRegression
from .mcmc import MCMCKernel


class TaylorProxy:
    """``HMCECS.taylor_proxy(reference_params, degree)``: reference values of
    the latent sites (constrained; the logistic model's sites are real)."""

    def __init__(self, reference_params, degree=2):
        if degree not in (1, 2):
            raise ValueError("Taylor proxy only defined for first and second degree.")
        self.reference_params = dict(reference_params)
        self.degree = int(degree)


def taylor_proxy(reference_params, degree=2):
    return TaylorProxy(reference_params, degree)


def _split_keys(keys, num, part):
    """random.split of each row of keys [K, 2] (uint32, cuda int32 bits)."""
    out = torch.empty((keys.shape[0], num, 2), dtype=torch.int32, device="cuda")
    _lib.check(_lib.load().nax_split_keys(_lib.ptr(keys), keys.shape[0], num, int(part),
                                          _lib.ptr(out)), "nax_split_keys")
    return out


class HMCECS(MCMCKernel):
    """``hmc_gibbs.py:502-689``.  ``inner_kernel`` is an ``HMC`` or ``NUTS``
    built on a model function; ``proxy`` is ``HMCECS.taylor_proxy(...)`` or
    None (the plate's n/m scaling of the subsample log-likelihood)."""

    def __init__(self, inner_kernel, *, num_blocks=1, proxy=None):
        if not is_model_function(getattr(inner_kernel, "_model", None)):
            raise ValueError("HMCECS needs an inner HMC/NUTS kernel built on a model function")
        if proxy is not None and not isinstance(proxy, TaylorProxy):
            raise ValueError("proxy must be HMCECS.taylor_proxy(...) or None")
        self.inner_kernel = inner_kernel
        self._num_blocks = int(num_blocks)
        self._proxy = proxy
        self._model = inner_kernel._model
        self._ecs_model = None
        self._gibbs_sites = ()

    taylor_proxy = staticmethod(taylor_proxy)

    @property
    def model(self):
        return self._model

    @property
    def sample_field(self):
        return "z"

    @property
    def default_fields(self):
        return ("z",)

    @property
    def _engine(self):
        return self.inner_kernel._engine

    def get_diagnostics_str(self, state):
        return self.inner_kernel.get_diagnostics_str(state)

    def postprocess_fn(self, args, kwargs):
        return self.inner_kernel.postprocess_fn(args, kwargs)

    def init(self, rng_key, num_warmup, init_params=None, model_args=(), model_kwargs=None):
        """``hmc_gibbs.py:577-638`` and ``123-152``.  Per chain key K:
        (rng_key, key_u) = split(K); the initial subsample is randint over the
        rows with split(key_u)[1] (drawn with replacement); with a proxy
        rng_key = split(rng_key)[0]; (rng_key, key_z) = split(rng_key); key_z
        initialises the inner kernel and rng_key is the HMCECS state key."""
        model_kwargs = model_kwargs or {}
        target = compile_model(self._model, *model_args, **model_kwargs)
        sub = target.params.get("subsample")
        if sub is None:
            raise AssertionError("Cannot detect any subsample statements in the model.")
        pname, n, m = sub
        self._gibbs_sites = (pname,)
        ref = None
        if self._proxy is not None:
            ref = target.flatten(self._proxy.reference_params).reshape(-1)
        ecs = SubsampledLogisticRegression(
            target.params["X"], target.params["y"], m, prior_scale=target.params["prior_scale"],
            num_blocks=self._num_blocks, proxy_degree=self._proxy.degree if self._proxy else 0,
            reference=ref, site_sizes=target.site_sizes)
        keys = np.asarray(rng_key, dtype=np.uint32).reshape(-1, 2)
        part = nrandom.PARTITIONABLE
        kd = torch.as_tensor(keys.view(np.int32)).to("cuda").contiguous()
        s1 = _split_keys(kd, 2, part)
        state_key, key_u = s1[:, 0].contiguous(), s1[:, 1].contiguous()
        key_idx = _split_keys(key_u, 2, part)[:, 1].contiguous()
        if self._proxy is not None:
            state_key = _split_keys(state_key, 2, part)[:, 0].contiguous()
        s3 = _split_keys(state_key, 2, part)
        state_key, key_z = s3[:, 0].contiguous(), s3[:, 1].contiguous()
        idx = torch.empty((keys.shape[0], m), dtype=torch.int32, device="cuda")
        _lib.check(_lib.load().nax_randint(_lib.ptr(key_idx), keys.shape[0], m, 0, n, int(part),
                                           _lib.ptr(idx)), "nax_randint")
        ecs.idx, ecs.keys = idx, state_key
        self._ecs_model = ecs
        if isinstance(init_params, dict):
            init_params = target.flatten(init_params)
        inner_keys = key_z.cpu().numpy().view(np.uint32)
        saved = self.inner_kernel._model
        self.inner_kernel._model = ecs
        try:
            return self.inner_kernel.init(inner_keys, num_warmup, init_params, (), {})
        finally:
            self.inner_kernel._model = saved

    @property
    def subsample_indices(self):
        """The current subsample of every chain, {plate name: [C, m] int32}."""
        return {self._gibbs_sites[0]: self._ecs_model.idx.clone()}

    @property
    def gibbs_accept_prob(self):
        """Acceptance probability of the last Gibbs update of each chain."""
        return self._engine.gibbs_accept.clone()

    def sample(self, state, model_args, model_kwargs):
        return self.inner_kernel.sample(state, model_args, model_kwargs)
