"""``sample`` / ``plate`` / ``deterministic`` and the two effect handlers the
model compiler needs (``trace``, ``substitute``).

Semantics follow ``numpyro/primitives.py``: ``sample`` (``:142-252``) sends a
message through the handler stack innermost first (``apply_stack``,
``:37-66``), an observed site returns ``obs``; ``plate`` (``:474-637``)
broadcasts the batch shape of every site inside it at its dim (the first
plate entered takes dim -1 unless ``dim`` is given, ``:556-590``);
``deterministic`` (``:318-341``) records a named value.  A latent site needs
a handler that provides its value: this library never draws from the prior
in Python (the engine draws its initial points itself), so a model runs under
``compile_model``'s tracer or under ``substitute``.
"""
from collections import OrderedDict

import torch

_STACK = []


class Messenger:
    def __init__(self, fn=None):
        self.fn = fn

    def __enter__(self):
        _STACK.append(self)
        return self

    def __exit__(self, *exc):
        assert _STACK and _STACK[-1] is self, "handler stack out of order"
        _STACK.pop()

    def process_message(self, msg):
        pass

    def postprocess_message(self, msg):
        pass

    def __call__(self, *args, **kwargs):
        with self:
            return self.fn(*args, **kwargs)


def apply_stack(msg):
    for h in reversed(_STACK):
        h.process_message(msg)
        if msg.get("stop"):
            break
    if msg["value"] is None and msg["type"] == "sample":
        raise RuntimeError(
            f"sample site {msg['name']!r} has no value: run the model through "
            "numpyro_amd.compile_model (or MCMC), or under substitute(...)")
    for h in _STACK:
        h.postprocess_message(msg)
    return msg


def _site_shape(fn, plates, sample_shape):
    shape = list(fn.batch_shape)
    for p in plates:
        n = -p.dim
        if len(shape) < n:
            shape = [1] * (n - len(shape)) + shape
        if shape[-n] not in (1, p.size):
            raise ValueError(f"plate {p.name!r} of size {p.size} does not match batch shape "
                             f"{tuple(fn.batch_shape)}")
        shape[-n] = p.size
    return tuple(sample_shape) + tuple(shape) + tuple(fn.event_shape)


def sample(name, fn, obs=None, rng_key=None, sample_shape=(), infer=None, obs_mask=None):
    if obs_mask is not None:
        raise NotImplementedError("obs_mask (imputation) is not supported")
    plates = [h for h in _STACK if isinstance(h, plate)]
    msg = {
        "type": "sample", "name": name, "fn": fn, "value": obs, "is_observed": obs is not None,
        "plates": plates, "shape": _site_shape(fn, plates, sample_shape),
        "infer": infer or {}, "stop": False,
    }
    return apply_stack(msg)["value"]


def deterministic(name, value):
    msg = {"type": "deterministic", "name": name, "value": value, "plates": [], "stop": False}
    return apply_stack(msg)["value"]


class plate(Messenger):
    """A plate of ``size``; with ``subsample_size < size`` it is a subsample
    plate for HMCECS.  The model compiler traces such a plate over all of its
    rows (the returned indices are ``arange(size)``): the HMCECS kernels draw
    and update each chain's subsample on the GPU."""

    def __init__(self, name, size, subsample_size=None, dim=None):
        self.name = name
        self.size = int(size)
        self.subsample_size = self.size if subsample_size is None else int(subsample_size)
        if not 0 < self.subsample_size <= self.size:
            raise ValueError("subsample_size must be in [1, size]")
        if dim is not None and dim >= 0:
            raise ValueError("dim arg must be negative.")
        self.dim = dim
        super().__init__()

    def __enter__(self):
        if self.dim is None:
            used = {p.dim for p in _STACK if isinstance(p, plate)}
            d = -1
            while d in used:
                d -= 1
            self.dim = d
        super().__enter__()
        return torch.arange(self.size)


def subsample(data, event_dim):
    """``primitives.py:640-680``: ``data`` indexed by the enclosing subsample
    plates' indices, which are every row under the model compiler."""
    for p in _STACK:
        if isinstance(p, plate) and p.subsample_size < p.size:
            dim = p.dim - event_dim
            if data.shape[dim] != p.size:
                raise ValueError(f"subsample: data has {data.shape[dim]} rows at dim {dim}, "
                                 f"plate {p.name!r} has {p.size}")
    return data


class trace(Messenger):
    """Records every site in order; ``get_trace(*args)`` runs the model."""

    def __enter__(self):
        super().__enter__()
        self.trace = OrderedDict()
        return self.trace

    def postprocess_message(self, msg):
        if msg["name"] in self.trace:
            raise ValueError(f"duplicate site name {msg['name']!r}")
        self.trace[msg["name"]] = dict(msg)

    def get_trace(self, *args, **kwargs):
        self(*args, **kwargs)
        return self.trace


class substitute(Messenger):
    """Gives latent sites the values in ``data`` (``handlers.py:811-866``)."""

    def __init__(self, fn=None, data=None):
        self.data = data or {}
        super().__init__(fn)

    def process_message(self, msg):
        if msg["type"] == "sample" and not msg["is_observed"] and msg["name"] in self.data:
            msg["value"] = self.data[msg["name"]]
