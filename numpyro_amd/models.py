"""Built-in potentials, evaluated by the engine's HIP kernels.

A model object stands where the reference passes ``potential_fn`` (the
negative log joint in unconstrained space, ``numpyro/infer/util.py:333-358``)
and ``postprocess_fn`` (constrain, ``infer/util.py:663-835``).  Latents are a
flat vector in sorted site-name order, as ``ravel_pytree`` flattens a dict.
"""
import math
import warnings

import torch

from . import _lib


def _check_binary_obs(y):
    """BernoulliLogits' support is {0, 1}; with validation on (numpyro's
    default, ``distributions/distribution.py:67, 430-440``) other observed
    values warn and give log_prob = -inf, so no initial point is valid.
    Returns True when every label is 0 or 1."""
    ok = bool(((y == 0) | (y == 1)).all())
    if not ok:
        warnings.warn("Out-of-support values provided to log prob method. "
                      "The value argument should be within the support.", stacklevel=3)
    return ok


POSITIVE_TRANSFORMS = {"exp": 0, "softplus": 1}


def constrain(v, tf):
    """Constrained value of an unconstrained site: ExpTransform ("exp",
    transforms.py:606-652), SoftplusTransform ("softplus",
    transforms.py:1156-1175) or the identity (None)."""
    if tf == "exp":
        return torch.exp(v)
    if tf == "softplus":
        return torch.nn.functional.softplus(v)
    return v


def unconstrain(v, tf):
    """Inverse of ``constrain``; softplus^-1(x) = log(-expm1(-x)) + x
    (``_softplus_inv``, transforms.py:1152-1153)."""
    if tf == "exp":
        return torch.log(v)
    if tf == "softplus":
        return torch.log(-torch.expm1(-v)) + v
    return v


def _positive_kind(positive_transform):
    if positive_transform not in POSITIVE_TRANSFORMS:
        raise ValueError("positive_transform must be 'exp' (biject_to(positive)) or "
                         "'softplus' (biject_to(softplus_positive))")
    return positive_transform


class Model:
    """Base class: a potential U(z) over a flat latent of size ``dim``."""

    dim = 0
    site_sizes = ()  # ((name, size, transform), ...) in sorted name order

    def attach(self, handle):
        raise NotImplementedError

    def postprocess(self, z):
        """Constrained values per site for unconstrained z [..., dim]."""
        out, o = {}, 0
        for name, size, tf in self.site_sizes:
            v = z[..., o:o + size]
            if size == 1:
                v = v[..., 0]
            out[name] = constrain(v, tf)
            o += size
        return out

    def algorithmic_bytes(self, num_chains):
        return 0.0

    def algorithmic_flops(self, num_chains):
        return 0.0


class LogisticRegression(Model):
    """``coefs ~ Normal(0, prior_scale)``, ``obs ~ Bernoulli(logits=X @ coefs)``
    (``examples/covtype.py:66-71``, ``test/infer/test_mcmc.py:104-168``).

    ``X`` [N, D] and ``y`` [N] are moved to the GPU as float32.  ``x_dtype``
    picks the storage the gradient kernels read: ``"fp32"`` (plain),
    ``"bf16"`` (MFMA; by default one bf16 copy of X with bf16 residuals in
    X^T R, ``kernel="two-copy"`` adds an fp16 copy for fp16 residuals) or
    ``"fp8"`` (OCP e4m3fn bytes with a power-of-two scale per column, decoded
    exactly in the VALU tile kernel, fp32 arithmetic; the quantised design
    defines the model).  The fp8 MFMA is not used: both of its operands are
    8-bit, so the positions would need a three-term e4m3 split (as costly as
    the two bf16 products) and the residuals of X^T R would keep 3 mantissa
    bits."""

    def __init__(self, X, y, prior_scale=1.0, x_dtype="fp32", kernel="auto"):
        X = torch.as_tensor(X)
        y = torch.as_tensor(y)
        if X.ndim != 2 or y.ndim != 1 or y.shape[0] != X.shape[0]:
            raise ValueError("X must be [N, D] and y [N]")
        self.X = X.to(device="cuda", dtype=torch.float32).contiguous()
        self.y = y.to(device="cuda", dtype=torch.float32).contiguous()
        self.N, self.dim = self.X.shape
        self.valid_obs = _check_binary_obs(self.y)
        self.prior_scale = float(prior_scale)
        if x_dtype not in _lib.X_DTYPES:
            raise ValueError(f"unknown x_dtype {x_dtype!r}")
        self.x_dtype = x_dtype
        if kernel not in ("auto", "tile", "mfma", "chain", "single", "two-copy"):
            raise ValueError("kernel must be 'auto', 'tile', 'mfma', 'chain', 'single' or "
                             "'two-copy'")
        if kernel in ("single", "two-copy") and x_dtype != "bf16":
            raise ValueError(f"kernel {kernel!r} (a bf16 form) needs x_dtype='bf16'")
        if x_dtype == "bf16" and kernel in ("tile", "chain"):
            raise ValueError("bf16 storage runs on the matrix cores: kernel 'auto', 'single' "
                             "or 'two-copy'")
        if _canon(x_dtype) == "fp8" and kernel not in ("auto", "tile"):
            raise ValueError("fp8 storage runs on the VALU tile kernel: kernel 'auto' or 'tile'")
        self.kernel = kernel
        self.site_sizes = (("coefs", self.dim, None),)

    def attach(self, handle):
        lib = _lib.load()
        kern = {"auto": 0, "tile": 1, "mfma": 2, "chain": 3, "single": 4, "two-copy": 2}[self.kernel]
        _lib.check(lib.nax_model_logistic_ex(handle, _lib.ptr(self.X), _lib.ptr(self.y), self.N,
                                             self.dim, self.prior_scale,
                                             _lib.X_DTYPES[self.x_dtype], kern),
                   "logistic model")

    def postprocess(self, z):
        return {"coefs": z}

    def algorithmic_flops(self, num_chains):
        # X.B and X^T.R: 2 * 2 N D flops per chain
        return 4.0 * self.N * self.dim * num_chains

    def algorithmic_bytes(self, num_chains):
        elt = {"fp32": 4, "bf16": 2, "fp8": 1}[_canon(self.x_dtype)]
        return elt * self.N * self.dim + 4.0 * self.N


class SubsampledLogisticRegression(Model):
    """The logistic model with a subsample plate over its rows
    (``examples/covtype.py:66-71`` with ``subsample_size``), for ``HMCECS``
    (``numpyro/infer/hmc_gibbs.py:502-``).  Each chain evaluates its own
    subsample of ``subsample_size`` rows of X (fp32), with the plate's n/m
    scaling (``proxy_degree=0``) or the difference estimator around
    ``reference`` coefficients (``proxy_degree`` 1 or 2,
    ``numpyro/contrib/ecs_proxies.py:23-50, 95-250``).  ``idx`` [C, m] int32
    and ``keys`` [C, 2] (the HMCECS state keys, as int32 bit patterns) are
    device tensors set by ``HMCECS`` before the engine is created."""

    def __init__(self, X, y, subsample_size, prior_scale=1.0, num_blocks=1, proxy_degree=0,
                 reference=None, site_sizes=None):
        X = torch.as_tensor(X)
        y = torch.as_tensor(y)
        if X.ndim != 2 or y.ndim != 1 or y.shape[0] != X.shape[0]:
            raise ValueError("X must be [N, D] and y [N]")
        self.X = X.to(device="cuda", dtype=torch.float32).contiguous()
        self.y = y.to(device="cuda", dtype=torch.float32).contiguous()
        self.N, self.dim = self.X.shape
        self.valid_obs = _check_binary_obs(self.y)
        self.prior_scale = float(prior_scale)
        self.m = int(subsample_size)
        if not 0 < self.m <= min(self.N, 4096):
            raise ValueError("subsample_size must be in [1, min(N, 4096)]")
        self.num_blocks = int(num_blocks)
        if self.num_blocks < 1:
            raise ValueError("num_blocks must be positive")
        self.proxy_degree = int(proxy_degree)
        if self.proxy_degree not in (0, 1, 2):
            raise ValueError("Taylor proxy only defined for first and second degree.")
        self.reference = None
        if self.proxy_degree:
            ref = torch.as_tensor(reference, dtype=torch.float64).reshape(-1)
            if ref.shape[0] != self.dim:
                raise ValueError(f"reference coefficients must have {self.dim} entries")
            self.reference = ref.to("cuda").contiguous()
        self.site_sizes = tuple(site_sizes) if site_sizes else (("coefs", self.dim, None),)
        self.idx = None
        self.keys = None

    def attach(self, handle):
        if self.idx is None or self.keys is None:
            raise RuntimeError("subsampled model: set idx and keys (HMCECS.init) first")
        _lib.check(_lib.load().nax_model_logistic_ecs(
            handle, _lib.ptr(self.X), _lib.ptr(self.y), self.N, self.dim, self.prior_scale,
            self.m, self.num_blocks, self.proxy_degree, _lib.ptr(self.reference),
            _lib.ptr(self.idx), _lib.ptr(self.keys)), "subsampled logistic model")

    def algorithmic_flops(self, num_chains):
        return 4.0 * self.m * self.dim * num_chains

    def algorithmic_bytes(self, num_chains):
        # per chain: each subsample row once, its label and index, and the
        # four fp64 proxy scalars of the row
        return num_chains * self.m * (4.0 * self.dim + 8.0 + (32.0 if self.proxy_degree else 0))


def _canon(x_dtype):
    return {0: "fp32", 1: "bf16", 2: "fp8"}[_lib.X_DTYPES[x_dtype]]


class StochasticVolatility(Model):
    """``examples/stochastic_volatility.py:58-66``: sigma ~ Exponential(50),
    s ~ GaussianRandomWalk(sigma, T), nu ~ Exponential(0.1) (other rates via
    ``sigma_rate`` / ``nu_rate``),
    r ~ StudentT(nu, 0, exp(s)).  Latent (sorted site names, positive sites
    through ExpTransform): ``[log nu, s (T), log sigma]``; with
    ``noncentred=True`` ``[log nu, log sigma, zeta (T)]``, s = sigma cumsum(zeta).
    ``positive_transform="softplus"`` unconstrains sigma and nu by
    SoftplusTransform instead (biject_to(softplus_positive))."""

    def __init__(self, returns, noncentred=False, sigma_rate=50.0, nu_rate=0.1,
                 positive_transform="exp"):
        r = torch.as_tensor(returns)
        if r.ndim != 1:
            raise ValueError("returns must be a vector")
        self.r = r.to(device="cuda", dtype=torch.float32).contiguous()
        self.T = int(self.r.shape[0])
        self.dim = self.T + 2
        self.noncentred = bool(noncentred)
        self.sigma_rate, self.nu_rate = float(sigma_rate), float(nu_rate)
        if not (self.sigma_rate > 0 and self.nu_rate > 0 and math.isfinite(self.sigma_rate)
                and math.isfinite(self.nu_rate)):
            raise ValueError("sigma_rate and nu_rate must be finite and positive")
        self.positive_transform = tf = _positive_kind(positive_transform)
        # execution order of the sample statements (stochastic_volatility.py:58-66)
        self.site_order = ("sigma", "zeta" if self.noncentred else "s", "nu")
        if self.noncentred:
            self.site_sizes = (("nu", 1, tf), ("sigma", 1, tf), ("zeta", self.T, None))
        else:
            self.site_sizes = (("nu", 1, tf), ("s", self.T, None), ("sigma", 1, tf))

    def attach(self, handle):
        lib = _lib.load()
        _lib.check(lib.nax_model_sv_ex(handle, _lib.ptr(self.r), self.T, int(self.noncentred),
                                       self.sigma_rate, self.nu_rate), "sv model")
        _lib.check(lib.nax_set_positive_transform(
            handle, POSITIVE_TRANSFORMS[self.positive_transform]), "positive transform")

    def postprocess(self, z):
        out = super().postprocess(z)
        if self.noncentred:
            out["s"] = out["sigma"][..., None] * torch.cumsum(out["zeta"], dim=-1)
        return out

    def algorithmic_bytes(self, num_chains):
        # per gradient: z in, dU out, returns in
        return 4.0 * num_chains * (2 * self.dim + self.T)


class HierarchicalLogistic(Model):
    """Independent non-centred hierarchical logistic regressions of uneven
    size (N_i rows, D_i <= 64 columns), ``chains_per_group`` chains each:
    eta ~ Normal(0,1)^D_i, mu ~ Normal(0,1), tau ~ HalfCauchy(1),
    beta = mu + tau eta, y ~ Bernoulli(logits = X_i beta).  A chain's latent
    (sorted site names) is ``[eta (D_i), mu, log tau]``, zero-padded to
    ``dim = max D_i + 2``; chain c belongs to group c // chains_per_group.
    ``positive_transform="softplus"`` unconstrains tau by SoftplusTransform
    (the latent then holds softplus^-1(tau) instead of log tau)."""

    def __init__(self, Xs, ys, chains_per_group=4, positive_transform="exp"):
        if len(Xs) != len(ys) or not Xs:
            raise ValueError("Xs and ys must be equally long, non-empty lists")
        xs = [torch.as_tensor(x, dtype=torch.float32) for x in Xs]
        yv = [torch.as_tensor(v, dtype=torch.float32) for v in ys]
        self.dims = [int(x.shape[1]) for x in xs]
        self.rows = [int(x.shape[0]) for x in xs]
        if any(d < 1 or d > 64 for d in self.dims):
            raise ValueError("every group needs 1 <= D_i <= 64")
        self.G = len(xs)
        self.K = int(chains_per_group)
        self.dim = max(self.dims) + 2
        self.X = torch.cat([x.reshape(-1) for x in xs]).to("cuda").contiguous()
        self.y = torch.cat(yv).to("cuda").contiguous()
        self.valid_obs = _check_binary_obs(self.y)
        elt = [0]
        for n, d in zip(self.rows, self.dims):
            elt.append(elt[-1] + n * d)
        row = [0]
        for n in self.rows:
            row.append(row[-1] + n)
        self.elt_off = torch.tensor(elt[:-1], dtype=torch.int64, device="cuda")
        self.row_off = torch.tensor(row, dtype=torch.int64, device="cuda")
        self.dims_t = torch.tensor(self.dims, dtype=torch.int32, device="cuda")
        self.site_sizes = ()
        self.positive_transform = _positive_kind(positive_transform)

    def attach(self, handle):
        lib = _lib.load()
        _lib.check(lib.nax_model_hier(
            handle, _lib.ptr(self.X), _lib.ptr(self.y), _lib.ptr(self.row_off),
            _lib.ptr(self.elt_off), _lib.ptr(self.dims_t), self.G, self.K), "hierarchical model")
        _lib.check(lib.nax_set_positive_transform(
            handle, POSITIVE_TRANSFORMS[self.positive_transform]), "positive transform")

    def postprocess(self, z):
        """{"eta": [..., max D] (zero past D_i), "mu", "tau"} per chain."""
        C = z.shape[0]
        dims = torch.tensor(self.dims, device=z.device).repeat_interleave(self.K)[:C]
        idx = dims.view((C,) + (1,) * (z.ndim - 1))
        mu = torch.gather(z, -1, idx.expand(z.shape[:-1] + (1,))).squeeze(-1)
        ut = torch.gather(z, -1, (idx + 1).expand(z.shape[:-1] + (1,))).squeeze(-1)
        ar = torch.arange(z.shape[-1], device=z.device)
        eta = torch.where(ar < idx, z, torch.zeros_like(z))[..., :max(self.dims)]
        return {"eta": eta, "mu": mu, "tau": constrain(ut, self.positive_transform)}

    def algorithmic_bytes(self, num_chains):
        return 4.0 * sum(n * d + n for n, d in zip(self.rows, self.dims))

    def algorithmic_flops(self, num_chains):
        return 4.0 * self.K * sum(n * d for n, d in zip(self.rows, self.dims))


class PotentialFn(Model):
    """A torch ``potential_fn(z) -> U`` on one latent vector [D]; the engine
    evaluates it for all chains with ``torch.func.vmap`` + ``grad``.  This is a
    test and convenience path: the user function runs as torch ops between
    engine steps."""

    def __init__(self, fn, dim):
        self.fn = fn
        self.dim = int(dim)
        self.site_sizes = ()

    def attach(self, handle):
        _lib.check(_lib.load().nax_model_external(handle), "external model")

    def postprocess(self, z):
        return z

    def value_and_grad(self, z):
        z64 = z.detach().to(torch.float64)

        def single(v):
            return self.fn(v)

        g, u = torch.func.vmap(torch.func.grad_and_value(single))(z64)
        return u, g


def normal_log_norm(scale):
    return math.log(math.sqrt(2 * math.pi) * scale)


# ---- the adjoint kernels behind torch.autograd ---------------------------------

def _device_value_and_grad(model, z):
    """U [C] (float64) and dU/dz [C, dim] (float32) of a built-in model at
    z [C, dim] from the library's fused potential + adjoint kernels."""
    from .infer.hmc import Engine  # local: infer imports this module
    C = int(z.shape[0])
    if hasattr(model, "bind"):
        model.bind(C)
    cache = model.__dict__.setdefault("_vg_engines", {})
    eng = cache.get(C)
    if eng is None:
        eng = cache[C] = Engine(model, C, "NUTS", 0)
    ld = _lib.stride(eng.h)
    zp = torch.zeros((C, ld), dtype=torch.float32, device="cuda")
    zp[:, :model.dim] = z.detach().to(device="cuda", dtype=torch.float32)
    pe = torch.empty(C, dtype=torch.float64, device="cuda")
    g = torch.zeros((C, ld), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    _lib.check(_lib.load().nax_potential_grad(eng.h, _lib.ptr(zp), ld, C, _lib.ptr(pe),
                                              _lib.ptr(g)), "nax_potential_grad")
    return pe, g[:, :model.dim].contiguous()


class _PotentialEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, z, model):
        pe, g = _device_value_and_grad(model, z)
        ctx.save_for_backward(g)
        return pe

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return (grad_out.to(g.dtype)[:, None] * g).to(torch.float32), None


def potential_energy(model, z):
    """``potential_energy`` (``numpyro/infer/util.py:333-358``) of a built-in
    model for a batch of unconstrained latents z [C, dim], as a differentiable
    torch function: the forward pass and its gradient both come from the
    model's hand-written HIP potential + adjoint kernel (one launch computes
    both; backward only scales the saved gradient).  Returns U [C] float64."""
    if isinstance(model, PotentialFn):
        raise ValueError("potential_energy needs a built-in model (HIP kernels)")
    z = torch.as_tensor(z)
    if z.ndim != 2 or z.shape[1] != model.dim:
        raise ValueError(f"z must be [C, {model.dim}]")
    return _PotentialEnergy.apply(z, model)
