"""MI355X-native HMC/NUTS engine with numpyro.infer's kernel API.

    from numpyro_amd import random
    from numpyro_amd.infer import MCMC, NUTS
    from numpyro_amd.models import LogisticRegression

or, with a model function compiled onto the HIP kernels,

    import numpyro_amd as npa
    from numpyro_amd import distributions as dist

    def model(X, y):
        b = npa.sample("b", dist.Normal(0.0, 1.0).expand([X.shape[1]]))
        npa.sample("y", dist.Bernoulli(logits=X @ b), obs=y)

    MCMC(NUTS(model), num_warmup=500, num_samples=500).run(key, X, y)
"""
__version__ = "0.1.0"

from . import diagnostics, distributions, random  # noqa: F401
from .primitives import deterministic, plate, sample, subsample  # noqa: F401
from . import compile  # noqa: F401,E402
