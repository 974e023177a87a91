"""MI355X-native HMC/NUTS engine with numpyro.infer's kernel API.

    from numpyro_amd import random
    from numpyro_amd.infer import MCMC, NUTS
    from numpyro_amd.models import LogisticRegression
"""
__version__ = "0.1.0"

from . import diagnostics, random  # noqa: F401
