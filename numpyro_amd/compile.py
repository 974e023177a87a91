"""Model -> kernel compiler for ``NUTS(model)``.

The reference turns a model function into ``potential_fn`` by tracing it
with effect handlers and summing log-densities in JAX
(``numpyro/infer/util.py:511-564`` ``initialize_model``,
``:300-358`` ``potential_energy``, ``:663-835`` ``postprocess``).  Here the
model is traced ONCE, symbolically: every latent site returns a ``Sym`` node,
torch operations on it are recorded through ``__torch_function__``, and the
resulting site graph is matched against the model families whose
log-density and adjoint exist as fused HIP kernels:

* logistic regression: ``b ~ Normal(0, s)`` [D] (any batch/event split of the
  D entries), ``y ~ Bernoulli(logits = X @ b)`` (or ``sum(b * X, -1)``,
  ``b @ X.T``), optionally ``+ a`` with an intercept ``a ~ Normal(0, s)``
  [N] -> ``models.LogisticRegression`` (the intercept as a column of ones) (``examples/covtype.py:66-71``,
  ``test/infer/test_mcmc.py:104-168``);
* stochastic volatility: ``sigma ~ Exponential(a)``, ``s ~
  GaussianRandomWalk(sigma, T)``, ``nu ~ Exponential(b)``, ``r ~ StudentT(nu,
  0, exp(s))`` -> ``models.StochasticVolatility``
  (``examples/stochastic_volatility.py:58-66``); the non-centred form
  ``zeta ~ Normal(0, 1)`` [T], ``s = sigma * cumsum(zeta)`` likewise;
* non-centred hierarchical logistic regression (one group): ``eta ~
  Normal(0, 1)`` [D], ``mu ~ Normal(0, 1)``, ``tau ~ HalfCauchy(1)``, ``y ~
  Bernoulli(logits = X @ (mu + tau * eta))`` -> ``models.HierarchicalLogistic``.

A model outside these forms is refused with ``NotImplementedError`` naming
the forms: there is no generic (non-kernel) path.  ``match_model`` returns
the matched form and its parameters without touching the GPU;
``compile_model`` builds the engine model object.
"""
import numbers

import numpy as np
import torch

from . import distributions as dist
from .models import (HierarchicalLogistic, LogisticRegression, Model, StochasticVolatility,
                     constrain)
from .primitives import Messenger, substitute, trace

# ---- symbolic values ----------------------------------------------------------


class Sym:
    """A value of the traced model: an op node over constants and latents.

    ``op`` is "latent" (args: site name), "const" (args: tensor) or a
    canonical operation name ("matmul", "mul", "add", "sub", "neg", "exp",
    "sum", "cumsum", "getitem", "unsqueeze", "reshape", "div").  ``shape`` is
    inferred by replaying the operation on meta tensors."""

    __slots__ = ("op", "args", "kwargs", "shape", "func")

    def __init__(self, op, args, shape, kwargs=None, func=None):
        self.op = op
        self.args = tuple(args)
        self.kwargs = dict(kwargs or {})
        self.shape = torch.Size(shape)
        self.func = func

    @property
    def ndim(self):
        return len(self.shape)

    def dim(self):
        return len(self.shape)

    def size(self, d=None):
        return self.shape if d is None else self.shape[d]

    def __len__(self):
        return self.shape[0]

    def __repr__(self):
        if self.op == "latent":
            return f"Sym(latent {self.args[0]!r}, {tuple(self.shape)})"
        return f"Sym({self.op}, {tuple(self.shape)})"

    @classmethod
    def __torch_function__(cls, func, types, args=(), kwargs=None):
        return _record(func, args, kwargs or {})

    # Python operators (torch dispatches the reflected forms to
    # __torch_function__ when the left operand is a tensor)
    def __add__(self, o):
        return _record(torch.add, (self, o), {})

    def __radd__(self, o):
        return _record(torch.add, (o, self), {})

    def __sub__(self, o):
        return _record(torch.sub, (self, o), {})

    def __rsub__(self, o):
        return _record(torch.sub, (o, self), {})

    def __mul__(self, o):
        return _record(torch.mul, (self, o), {})

    def __rmul__(self, o):
        return _record(torch.mul, (o, self), {})

    def __truediv__(self, o):
        return _record(torch.div, (self, o), {})

    def __neg__(self):
        return _record(torch.neg, (self,), {})

    def __matmul__(self, o):
        return _record(torch.matmul, (self, o), {})

    def __rmatmul__(self, o):
        return _record(torch.matmul, (o, self), {})

    def __getitem__(self, idx):
        return _record(torch.Tensor.__getitem__, (self, idx), {})

    def exp(self):
        return _record(torch.exp, (self,), {})

    def sum(self, *a, **k):
        return _record(torch.sum, (self,) + a, k)

    def cumsum(self, *a, **k):
        return _record(torch.cumsum, (self,) + a, k)

    def unsqueeze(self, d):
        return _record(torch.unsqueeze, (self, d), {})

    def reshape(self, *shape):
        return _record(torch.reshape, (self, shape[0] if len(shape) == 1 else shape), {})


def _canon(func):
    name = getattr(func, "__name__", str(func)).strip("_")
    table = {
        "matmul": "matmul", "rmatmul": "rmatmul", "mv": "matmul", "mm": "matmul",
        "mul": "mul", "rmul": "mul", "multiply": "mul",
        "add": "add", "radd": "add", "sub": "sub", "rsub": "rsub", "subtract": "sub",
        "neg": "neg", "negative": "neg", "exp": "exp", "sum": "sum", "cumsum": "cumsum",
        "getitem": "getitem", "unsqueeze": "unsqueeze", "reshape": "reshape", "view": "reshape",
        "div": "div", "truediv": "div", "true_divide": "div",
    }
    return table.get(name)


def _meta(v):
    if isinstance(v, Sym):
        if v.op == "const":
            t = v.args[0]
            return torch.zeros(t.shape, dtype=t.dtype, device="meta")
        return torch.zeros(v.shape, device="meta")
    if torch.is_tensor(v):
        return torch.zeros(v.shape, dtype=v.dtype, device="meta")
    if isinstance(v, (tuple, list)):
        return type(v)(_meta(x) for x in v)
    return v


def _const(v):
    if isinstance(v, Sym):
        return v
    if isinstance(v, np.ndarray):
        v = torch.as_tensor(v)
    if torch.is_tensor(v):
        return Sym("const", (v,), v.shape)
    if isinstance(v, numbers.Number) and not isinstance(v, bool):
        return Sym("const", (torch.tensor(float(v), dtype=torch.float64),), ())
    raise NotImplementedError(f"unsupported operand {type(v).__name__} in a traced operation")


# elementwise / linear operations: every operand becomes a Sym
_ARITH = {"add": torch.add, "sub": torch.sub, "mul": torch.mul, "div": torch.div,
          "matmul": torch.matmul}


def _record(func, args, kwargs):
    op = _canon(func)
    if op is None:
        raise NotImplementedError(
            f"operation {getattr(func, '__name__', func)!r} on a latent value is not supported "
            "by the model compiler")
    if op == "rmatmul":  # Tensor.__rmatmul__(self, other) is other @ self
        op, args = "matmul", (args[1], args[0])
    if op == "rsub":  # Tensor.__rsub__(self, other) is other - self
        op, args = "sub", (args[1], args[0])
    if kwargs.get("alpha", 1) != 1:
        raise NotImplementedError("add/sub with alpha is not supported")
    if op in _ARITH:
        if len(args) != 2 or kwargs:
            raise NotImplementedError(f"{op} with extra arguments is not supported")
        args = tuple(_const(a) for a in args)
        func = _ARITH[op]
    try:
        shape = _apply(func, [_meta(a) for a in args],
                       {k: _meta(v) for k, v in kwargs.items()}).shape
    except Exception as e:  # noqa: BLE001 - a shape error is the model's own bug
        raise ValueError(f"shape error in traced operation {op}: {e}") from e
    return Sym(op, args, shape, kwargs, func)


def _apply(func, args, kwargs):
    with torch._C.DisableTorchFunction():
        return func(*args, **kwargs)


# ---- tracing ------------------------------------------------------------------


class _symbolic(Messenger):
    def process_message(self, msg):
        if msg["type"] == "sample" and not msg["is_observed"]:
            fn = msg["fn"]
            if fn.base.support not in (dist.REAL, dist.POSITIVE):
                raise NotImplementedError(
                    f"latent site {msg['name']!r}: only continuous latent sites are supported")
            msg["value"] = Sym("latent", (msg["name"],), msg["shape"])


def trace_model(model, *args, **kwargs):
    """The model's sites, in execution order, with symbolic latent values."""
    with trace() as tr, _symbolic():
        model(*args, **kwargs)
    return tr


# ---- pattern matching ---------------------------------------------------------


def _is_const(v):
    return isinstance(v, Sym) and v.op == "const" or torch.is_tensor(v) or isinstance(
        v, (numbers.Number, np.ndarray))


def _cval(v):
    if isinstance(v, Sym):
        if v.op != "const":
            raise _NoMatch
        return v.args[0]
    if isinstance(v, numbers.Number):
        return torch.tensor(float(v), dtype=torch.float64)
    return torch.as_tensor(v)


def _uniform_scalar(v, want=None):
    """The single value a constant holds everywhere; no match if it is not a
    constant, varies, or differs from ``want``."""
    if not _is_const(v):
        raise _NoMatch
    t = _cval(v).to(torch.float64).reshape(-1)
    if t.numel() == 0:
        raise _NoMatch
    x = float(t[0])
    if not bool((t == x).all()):
        raise _NoMatch
    if want is not None and x != want:
        raise _NoMatch
    return x


class _NoMatch(Exception):
    pass


def _latent(v):
    if isinstance(v, Sym) and v.op == "latent":
        return v.args[0]
    raise _NoMatch


def _linear(v):
    """v = X @ b (X constant [N, D]) or sum(b * X, -1) / sum(X * b, -1);
    returns (X, b-expression)."""
    if not isinstance(v, Sym):
        raise _NoMatch
    if v.op == "matmul":
        a, b = v.args
        if _is_const(a) and isinstance(b, Sym) and b.op != "const":
            X = _cval(a)
            if X.ndim == 2 and len(b.shape) == 1:
                return X, b
        if _is_const(b) and isinstance(a, Sym) and a.op != "const":  # b @ X^T
            Xt = _cval(b)
            if Xt.ndim == 2 and len(a.shape) == 1:
                return Xt.T, a
        raise _NoMatch
    if v.op == "sum":
        dims = v.args[1:] if len(v.args) > 1 else (v.kwargs.get("dim", v.kwargs.get("axis")),)
        inner = v.args[0]
        if len(dims) != 1 or dims[0] not in (-1, 1) or v.kwargs.get("keepdim", False):
            raise _NoMatch
        if isinstance(inner, Sym) and inner.op == "mul" and len(inner.shape) == 2:
            a, b = inner.args
            if _is_const(a) and isinstance(b, Sym) and b.op != "const":
                X, e = _cval(a), b
            elif _is_const(b) and isinstance(a, Sym) and a.op != "const":
                X, e = _cval(b), a
            else:
                raise _NoMatch
            if X.ndim == 2 and len(e.shape) == 1 and tuple(X.shape) == tuple(inner.shape):
                return X, e
    raise _NoMatch


def _flat_size(shape):
    return int(np.prod(shape)) if len(shape) else 1


def _check_obs_shape(site, n):
    if _flat_size(site["shape"]) != n:
        raise _NoMatch


def _normal_scale(site, name):
    pd = site["fn"].base
    if not isinstance(pd, dist.Normal):
        raise _NoMatch
    _uniform_scalar(pd.loc, 0.0)
    s = _uniform_scalar(pd.scale)
    if not s > 0:
        raise ValueError(f"site {name!r}: Normal scale must be positive, got {s}")
    return s


def _match_logistic(lat, obs):
    """b ~ Normal(0, s) [D], optionally a ~ Normal(0, s) scalar intercept
    with the same s, y ~ Bernoulli(logits = X b [+ a]).  The intercept is a
    coefficient of the kernel on a column of ones, placed where the
    reference's flattening puts its site: ravel_pytree orders the latent dict
    by sorted site name, so the ones column comes first when the intercept's
    name sorts before the coefficients' and last otherwise."""
    if len(lat) not in (1, 2) or len(obs) != 1:
        raise _NoMatch
    (oname, osite), = obs.items()
    od = osite["fn"].base
    if not isinstance(od, dist.BernoulliLogits):
        raise _NoMatch
    logits, aname = od.logits, None
    if len(lat) == 2:
        if not (isinstance(logits, Sym) and logits.op == "add"):
            raise _NoMatch
        p, q = logits.args
        if isinstance(p, Sym) and p.op == "latent":
            p, q = q, p
        aname = _latent(q)
        if aname not in lat or _flat_size(lat[aname]["shape"]) != 1:
            raise _NoMatch
        logits = p
    X, e = _linear(logits)
    bname = _latent(e)
    if bname not in lat or bname == aname or len(lat[bname]["shape"]) != 1 or \
            X.shape[1] != lat[bname]["shape"][0]:
        raise _NoMatch
    s = _normal_scale(lat[bname], bname)
    _check_obs_shape(osite, X.shape[0])
    D = X.shape[1]
    sites = ((bname, D, None),)
    if aname is not None:
        if _normal_scale(lat[aname], aname) != s:
            raise _NoMatch  # one prior scale for all coefficients
        ones = torch.ones((X.shape[0], 1), dtype=X.dtype, device=X.device)
        if aname < bname:
            X = torch.cat([ones, X], 1)
            sites = ((aname, 1, None),) + sites
        else:
            X = torch.cat([X, ones], 1)
            sites = sites + ((aname, 1, None),)
    return "logistic", dict(X=X, y=_cval(osite["value"]).reshape(-1), prior_scale=s,
                            sites=sites)


def _match_sv(lat, obs):
    if len(lat) != 3 or len(obs) != 1:
        raise _NoMatch
    (oname, osite), = obs.items()
    od = osite["fn"].base
    if not isinstance(od, dist.StudentT):
        raise _NoMatch
    _uniform_scalar(od.loc, 0.0)
    nu = _latent(od.df)
    sc = od.scale
    if not (isinstance(sc, Sym) and sc.op == "exp"):
        raise _NoMatch
    s_expr = sc.args[0]
    r = _cval(osite["value"]).reshape(-1)
    T = r.shape[0]
    _check_obs_shape(osite, T)
    nu_d = lat[nu]["fn"].base
    if not isinstance(nu_d, dist.Exponential) or _flat_size(lat[nu]["shape"]) != 1:
        raise _NoMatch
    nu_rate = _uniform_scalar(nu_d.rate)
    if isinstance(s_expr, Sym) and s_expr.op == "latent":
        s = s_expr.args[0]
        grw = lat[s]["fn"].base
        if not isinstance(grw, dist.GaussianRandomWalk) or grw.num_steps != T or \
                tuple(lat[s]["shape"]) != (T,):
            raise _NoMatch
        sigma = _latent(grw.scale)
        noncentred, field = False, s
    else:
        # s = sigma * cumsum(zeta) (either operand order, cumsum along the
        # only axis)
        if not (isinstance(s_expr, Sym) and s_expr.op == "mul"):
            raise _NoMatch
        a, b = s_expr.args
        if isinstance(a, Sym) and a.op == "cumsum":
            a, b = b, a
        if not (isinstance(b, Sym) and b.op == "cumsum"):
            raise _NoMatch
        cdim = b.args[1] if len(b.args) > 1 else b.kwargs.get("dim")
        if cdim not in (0, -1):
            raise _NoMatch
        sigma = _latent(a)
        field = _latent(b.args[0])
        zd = lat[field]["fn"].base
        if not isinstance(zd, dist.Normal) or tuple(lat[field]["shape"]) != (T,):
            raise _NoMatch
        _uniform_scalar(zd.loc, 0.0)
        _uniform_scalar(zd.scale, 1.0)
        noncentred = True
    sd = lat[sigma]["fn"].base
    if not isinstance(sd, dist.Exponential) or _flat_size(lat[sigma]["shape"]) != 1:
        raise _NoMatch
    sigma_rate = _uniform_scalar(sd.rate)
    if len({nu, sigma, field}) != 3:
        raise _NoMatch
    order = [(nu, 1, "exp"), (field, T, None), (sigma, 1, "exp")] if not noncentred else \
        [(nu, 1, "exp"), (sigma, 1, "exp"), (field, T, None)]
    return "sv", dict(returns=r, noncentred=noncentred, sigma_rate=sigma_rate,
                      nu_rate=nu_rate, sites=tuple(order), names=dict(nu=nu, sigma=sigma,
                                                                      field=field))


def _match_hier(lat, obs):
    if len(lat) != 3 or len(obs) != 1:
        raise _NoMatch
    (oname, osite), = obs.items()
    od = osite["fn"].base
    if not isinstance(od, dist.BernoulliLogits):
        raise _NoMatch
    X, e = _linear(od.logits)
    # e = mu + tau * eta (operands in either order)
    if not (isinstance(e, Sym) and e.op == "add"):
        raise _NoMatch
    a, b = e.args
    if isinstance(a, Sym) and a.op == "mul":
        a, b = b, a
    if not (isinstance(b, Sym) and b.op == "mul"):
        raise _NoMatch
    mu = _latent(a)
    p, q = b.args
    if isinstance(p, Sym) and p.op == "latent" and len(p.shape) == 1 and p.shape[0] > 1:
        p, q = q, p
    tau, eta = _latent(p), _latent(q)
    D = X.shape[1]
    if len({mu, tau, eta}) != 3 or tuple(lat[eta]["shape"]) != (D,):
        raise _NoMatch
    for nm in (mu, tau):
        if _flat_size(lat[nm]["shape"]) != 1:
            raise _NoMatch
    ed, md, td = lat[eta]["fn"].base, lat[mu]["fn"].base, lat[tau]["fn"].base
    if not (isinstance(ed, dist.Normal) and isinstance(md, dist.Normal)
            and isinstance(td, dist.HalfCauchy)):
        raise _NoMatch
    for d_ in (ed, md):
        _uniform_scalar(d_.loc, 0.0)
        _uniform_scalar(d_.scale, 1.0)
    _uniform_scalar(td.scale, 1.0)
    if not 1 <= D <= 64:
        raise _NoMatch
    _check_obs_shape(osite, X.shape[0])
    return "hier", dict(X=X, y=_cval(osite["value"]).reshape(-1),
                        sites=((eta, D, None), (mu, 1, None), (tau, 1, "exp")))


_FORMS = (_match_logistic, _match_sv, _match_hier)

_REFUSAL = (
    "the model does not match a form with fused HIP kernels: logistic regression "
    "(b ~ Normal(0, s) [D], optional intercept a ~ Normal(0, s); "
    "y ~ Bernoulli(logits=X @ b [+ a])), stochastic volatility "
    "(Exponential sigma and nu, GaussianRandomWalk s or s = sigma * cumsum(zeta) with "
    "zeta ~ Normal(0, 1), StudentT(nu, 0, exp(s)) returns) or non-centred hierarchical "
    "logistic regression (eta ~ Normal(0, 1) [D <= 64], mu ~ Normal(0, 1), "
    "tau ~ HalfCauchy(1), y ~ Bernoulli(logits=X @ (mu + tau * eta)))")


def match_model(model, *args, **kwargs):
    """Traces ``model(*args, **kwargs)`` and returns ``(form, params, trace)``
    with form in {"logistic", "sv", "hier"}.  Deterministic sites are kept in
    the trace (evaluated later from the latent values)."""
    tr = trace_model(model, *args, **kwargs)
    lat = {k: v for k, v in tr.items() if v["type"] == "sample" and not v["is_observed"]}
    obs = {k: v for k, v in tr.items() if v["type"] == "sample" and v["is_observed"]}
    for k, v in obs.items():
        if isinstance(v["value"], Sym):
            raise NotImplementedError(f"observed site {k!r}: the observation must be data")
    sub = {p.name: (p.size, p.subsample_size) for v in obs.values() for p in v["plates"]
           if p.subsample_size < p.size}
    for form in _FORMS:
        try:
            kind, params = form(lat, obs)
        except _NoMatch:
            continue
        if sub:
            (pname, (n, m)), = sub.items() if len(sub) == 1 else _many_subsample_plates(sub)
            if kind != "logistic" or params["X"].shape[0] != n:
                raise NotImplementedError(
                    "subsample plates (HMCECS) are supported for logistic regression, the "
                    "plate running over the rows of X")
            params["subsample"] = (pname, n, m)
        return kind, params, tr
    raise NotImplementedError(_REFUSAL)


def _many_subsample_plates(sub):
    raise NotImplementedError(f"more than one subsample plate: {sorted(sub)}")


# ---- building the engine model ----------------------------------------------


def compile_model(model, *args, x_dtype="fp32", kernel="auto", **kwargs):
    """Compiles ``model(*args, **kwargs)`` into an engine model (a
    ``numpyro_amd.models.Model``) whose potential and gradient are the HIP
    kernels of the matched form.  ``x_dtype`` / ``kernel`` choose the logistic
    kernels as in ``LogisticRegression``.  The returned object's
    ``postprocess`` yields the model's own site names (constrained values),
    plus its deterministic sites."""
    kind, p, tr = match_model(model, *args, **kwargs)
    return CompiledModel(kind, p, model, args, kwargs, tr, x_dtype=x_dtype, kernel=kernel)


def _deterministic_names(tr):
    return [k for k, v in tr.items() if v["type"] == "deterministic"]


class CompiledModel(Model):
    """An engine model produced by ``compile_model``: the kernels are those of
    the matched built-in model; ``site_sizes`` lists the user's sites in the
    kernel's latent order (name, size, "exp" for positive sites)."""

    def __init__(self, kind, params, fn, args, kwargs, tr, x_dtype="fp32", kernel="auto"):
        self.kind = kind
        self.params = params
        self.site_sizes = tuple(params["sites"])
        self.dim = sum(size for _, size, _ in self.site_sizes)
        # latent sites in the order the model samples them (the order
        # find_valid_initial_params draws their initial values in)
        self.site_order = tuple(k for k, v in tr.items()
                                if v["type"] == "sample" and not v["is_observed"])
        self.fn, self.args, self.kwargs = fn, args, kwargs
        self.deterministic = _deterministic_names(tr)
        self.x_dtype, self.kernel = x_dtype, kernel
        self.inner = None
        self._chains = None
        if kind == "logistic":
            self.inner = LogisticRegression(params["X"], params["y"],
                                            prior_scale=params["prior_scale"], x_dtype=x_dtype,
                                            kernel=kernel)
        elif kind == "sv":
            self.inner = StochasticVolatility(params["returns"], noncentred=params["noncentred"],
                                              sigma_rate=params["sigma_rate"],
                                              nu_rate=params["nu_rate"])
        self.valid_obs = True if self.inner is None else getattr(self.inner, "valid_obs", True)

    def bind(self, num_chains):
        """Fixes the number of chains.  The hierarchical kernel runs groups of
        K chains on one model (K in {1, 2, 4, 8, 16}); C chains of this one
        model become C / K copies of it, K the largest such divisor of C."""
        if self.kind != "hier" or self._chains == num_chains:
            return self
        C = int(num_chains)
        K = max(k for k in (16, 8, 4, 2, 1) if C % k == 0)
        G = C // K
        X, y = self.params["X"], self.params["y"]
        self.inner = HierarchicalLogistic([X] * G, [y] * G, chains_per_group=K)
        self.valid_obs = self.inner.valid_obs
        self._chains = C
        return self

    def attach(self, handle):
        if self.inner is None:
            raise RuntimeError("compiled hierarchical model: call bind(num_chains) first")
        self.inner.attach(handle)

    def unflatten(self, z):
        """{site: constrained value} for unconstrained z [..., dim]."""
        out, o = {}, 0
        for name, size, tf in self.site_sizes:
            v = z[..., o:o + size]
            if size == 1:
                v = v[..., 0]
            out[name] = constrain(v, tf)
            o += size
        return out

    def flatten(self, params):
        """Flat latent [..., dim] from {site: unconstrained value} (the form
        of ``init_params`` for a model, ``hmc.py:740-800``)."""
        parts = []
        for name, size, _ in self.site_sizes:
            v = torch.as_tensor(params[name], dtype=torch.float32)
            lead = v.shape[:v.ndim - 1] if size > 1 else v.shape
            parts.append(v.reshape(tuple(lead) + (size,)))
        return torch.cat(parts, -1)

    def postprocess(self, z):
        out = self.unflatten(z)
        if self.deterministic:
            out.update(self._deterministic_values(out))
        return out

    def _deterministic_values(self, values):
        """Replays the model under ``substitute`` for every draw (batched with
        torch.func.vmap over the leading dimensions)."""
        lead = next(iter(values.values())).shape
        nlead = len(lead) - (1 if self.site_sizes[0][1] > 1 else 0)
        lead = lead[:nlead]
        flat = {k: v.reshape((-1,) + tuple(v.shape[nlead:])) for k, v in values.items()}

        def one(vals):
            with trace() as tr, substitute(data=vals):
                self.fn(*self.args, **self.kwargs)
            return {k: tr[k]["value"] for k in self.deterministic}

        res = torch.func.vmap(one)(flat)
        return {k: v.reshape(tuple(lead) + tuple(v.shape[1:])) for k, v in res.items()}

    def algorithmic_bytes(self, num_chains):
        return self.inner.algorithmic_bytes(num_chains) if self.inner else 0.0

    def algorithmic_flops(self, num_chains):
        return self.inner.algorithmic_flops(num_chains) if self.inner else 0.0


def is_model_function(obj):
    return callable(obj) and not isinstance(obj, Model)


__all__ = ["Sym", "trace_model", "match_model", "compile_model", "CompiledModel"]
