"""Initialisation strategies for model-based kernels
(``numpyro/infer/initialization.py``).

The default draws every unconstrained latent coordinate uniformly from
(-2, 2) (``init_to_uniform``, ``initialization.py:95-129``).
``init_to_value(values=...)`` starts the sites named in ``values`` at those
(constrained) values and the others as the default does
(``initialization.py:140-156``); ``init_to_feasible`` starts every
unconstrained coordinate at 0 (``initialization.py:132-137``).
"""
import torch

from ..models import unconstrain


class init_to_value:
    """``init_to_value(site=None, values={})``."""

    def __init__(self, site=None, values=None):
        if site is not None:
            raise ValueError("init_to_value is applied by the kernel: pass values= only")
        self.values = dict(values or {})

    def unconstrained(self, site_sizes):
        """{site: unconstrained float32 values [1 or C, size]} of the named sites
        (their transform's inverse applied: log for ExpTransform sites,
        softplus^-1 for SoftplusTransform sites), as initialize_model hands
        them to find_valid_initial_params (numpyro/infer/util.py:776-779)."""
        names = {name for name, _, _ in site_sizes}
        unknown = set(self.values) - names
        if unknown:
            raise ValueError(f"init_to_value: no latent site named {sorted(unknown)}")
        out = {}
        for name, size, tf in site_sizes:
            if name in self.values:
                v = unconstrain(torch.as_tensor(self.values[name], dtype=torch.float64).cpu(),
                                tf)
                out[name] = v.reshape(-1, size).to(torch.float32).numpy()  # [1 or C, size]
        return out

    def apply(self, z0, site_sizes):
        """Overwrites the coordinates of the named sites in z0 [C, dim]."""
        z0 = torch.as_tensor(z0, dtype=torch.float32).clone()
        vals = self.unconstrained(site_sizes)
        o = 0
        for name, size, _ in site_sizes:
            if name in vals:
                z0[:, o:o + size] = torch.from_numpy(vals[name])
            o += size
        return z0


class init_to_feasible:
    """``init_to_feasible(site=None)``: every unconstrained coordinate 0."""

    def __init__(self, site=None):
        if site is not None:
            raise ValueError("init_to_feasible is applied by the kernel")

    def unconstrained(self, site_sizes):
        return "feasible"

    def apply(self, z0, site_sizes):
        return torch.zeros_like(torch.as_tensor(z0, dtype=torch.float32))
