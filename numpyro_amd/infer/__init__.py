from .hmc_gibbs import HMCECS, taylor_proxy  # noqa: F401
from .hmc import HMC, NUTS, HMCAdaptState, HMCState, hmc, init_to_uniform  # noqa: F401
from .initialization import init_to_feasible, init_to_value  # noqa: F401
from .mcmc import MCMC, MCMCKernel  # noqa: F401
from .util import fori_collect  # noqa: F401
