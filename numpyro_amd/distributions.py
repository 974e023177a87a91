"""Distribution descriptors for models written with ``numpyro_amd.sample``.

A model function built from these is compiled (``numpyro_amd.compile``) into
one of the engine's model descriptors, whose log-densities and adjoints run
in the HIP kernels; the classes here carry parameters and shapes only.  The
constructors, shape rules (``batch_shape`` from broadcasting the parameters,
``expand``, ``to_event``) and supports follow ``numpyro/distributions``:
Normal ``continuous.py:2408``, HalfCauchy ``continuous.py:1180``,
Exponential ``continuous.py:697``, StudentT ``continuous.py:3190``,
GaussianRandomWalk ``continuous.py:1072``, Bernoulli / BernoulliLogits
``discrete.py:95-160``, ExpandedDistribution / Independent
``distribution.py:700-1040``.
"""
import torch

REAL = "real"
POSITIVE = "positive"
BOOLEAN = "boolean"


def _shape(v):
    s = getattr(v, "shape", None)
    return tuple(s) if s is not None else ()


class Distribution:
    """Parameters (tensors, numbers or traced values) plus shapes."""

    support = REAL
    arg_names = ()

    def __init__(self, batch_shape=(), event_shape=()):
        self.batch_shape = tuple(batch_shape)
        self.event_shape = tuple(event_shape)

    def params(self):
        return {k: getattr(self, k) for k in self.arg_names}

    @property
    def shape_full(self):
        return self.batch_shape + self.event_shape

    def expand(self, batch_shape):
        return ExpandedDistribution(self, tuple(batch_shape))

    def to_event(self, reinterpreted_batch_ndims=None):
        if reinterpreted_batch_ndims is None:
            reinterpreted_batch_ndims = len(self.batch_shape)
        if reinterpreted_batch_ndims == 0:
            return self
        return Independent(self, reinterpreted_batch_ndims)

    @property
    def base(self):
        return self

    def __repr__(self):
        return f"{type(self).__name__}({', '.join(f'{k}={v!r}' for k, v in self.params().items())})"


def _batch(*vals):
    return tuple(torch.broadcast_shapes(*[_shape(v) for v in vals]))


class Normal(Distribution):
    arg_names = ("loc", "scale")

    def __init__(self, loc=0.0, scale=1.0):
        self.loc, self.scale = loc, scale
        super().__init__(_batch(loc, scale))


class HalfCauchy(Distribution):
    arg_names = ("scale",)
    support = POSITIVE

    def __init__(self, scale=1.0):
        self.scale = scale
        super().__init__(_batch(scale))


class Exponential(Distribution):
    arg_names = ("rate",)
    support = POSITIVE

    def __init__(self, rate=1.0):
        self.rate = rate
        super().__init__(_batch(rate))


class StudentT(Distribution):
    arg_names = ("df", "loc", "scale")

    def __init__(self, df, loc=0.0, scale=1.0):
        self.df, self.loc, self.scale = df, loc, scale
        super().__init__(_batch(df, loc, scale))


class GaussianRandomWalk(Distribution):
    """x_0 ~ N(0, scale), x_t - x_{t-1} ~ N(0, scale); event shape (num_steps,)."""
    arg_names = ("scale", "num_steps")

    def __init__(self, scale=1.0, num_steps=1):
        if int(num_steps) < 1:
            raise ValueError("`num_steps` argument should be an positive integer.")
        self.scale, self.num_steps = scale, int(num_steps)
        super().__init__(_batch(scale), (self.num_steps,))


class BernoulliLogits(Distribution):
    arg_names = ("logits",)
    support = BOOLEAN

    def __init__(self, logits):
        self.logits = logits
        super().__init__(_batch(logits))


def Bernoulli(probs=None, logits=None):
    """``discrete.py:150-160``: exactly one of ``probs`` / ``logits``; only the
    logits form has HIP kernels."""
    if (probs is None) == (logits is None):
        raise ValueError("Either `probs` or `logits` must be specified, but not both.")
    if probs is not None:
        raise NotImplementedError("Bernoulli(probs=...) is not supported; use logits=")
    return BernoulliLogits(logits)


class ExpandedDistribution(Distribution):
    def __init__(self, base_dist, batch_shape):
        bs = tuple(base_dist.batch_shape)
        if len(batch_shape) < len(bs) or any(
                b != 1 and b != e for b, e in zip(bs[::-1], batch_shape[::-1])):
            raise ValueError(f"cannot expand batch shape {bs} to {batch_shape}")
        self.base_dist = base_dist
        self.support = base_dist.support
        super().__init__(batch_shape, base_dist.event_shape)

    def params(self):
        return self.base_dist.params()

    @property
    def base(self):
        return self.base_dist.base


class Independent(Distribution):
    def __init__(self, base_dist, reinterpreted_batch_ndims):
        n = int(reinterpreted_batch_ndims)
        if n > len(base_dist.batch_shape):
            raise ValueError("Expected reinterpreted_batch_ndims <= len(base_distribution.batch_shape)")
        self.base_dist = base_dist
        self.support = base_dist.support
        bs = base_dist.batch_shape
        super().__init__(bs[:len(bs) - n], bs[len(bs) - n:] + base_dist.event_shape)

    def params(self):
        return self.base_dist.params()

    @property
    def base(self):
        return self.base_dist.base
