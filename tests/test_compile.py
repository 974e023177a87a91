"""Model compiler (NUTS(model) -> kernel descriptor), on the CPU.

The mapping is pinned independently of the matcher: the model is re-run
under ``substitute`` with concrete values, and its log joint is summed site
by site from the oracle's scipy-pinned log-densities (plus log|J| of the exp
bijection for positive sites, numpyro/infer/util.py:300-358).  That sum must
equal the oracle potential of the matched form with the matched parameters,
evaluated on the flat latent laid out in the compiled site order.  A wrong
prior scale or rate, a swapped site or a missing Jacobian fails it.
"""
import numpy as np
import pytest
import torch

import numpyro_amd as npa
from numpyro_amd import distributions as dist
from numpyro_amd.compile import match_model
from numpyro_amd.primitives import substitute, trace
from oracle import distributions as od
from oracle import models as om


def _site_log_prob(site):
    """log p of one site from the oracle, with concrete parameters."""
    fn = site["fn"].base
    v = np.asarray(torch.as_tensor(site["value"]), dtype=np.float64)
    p = {k: np.asarray(torch.as_tensor(x), dtype=np.float64) if not isinstance(x, int) else x
         for k, x in fn.params().items()}
    if isinstance(fn, dist.Normal):
        lp = od.normal_log_prob(v, p["loc"], p["scale"])
    elif isinstance(fn, dist.HalfCauchy):
        lp = od.half_cauchy_log_prob(v, p["scale"])
    elif isinstance(fn, dist.Exponential):
        lp = od.exponential_log_prob(v, p["rate"])
    elif isinstance(fn, dist.StudentT):
        lp = od.student_t_log_prob(v, p["df"], p["loc"], p["scale"])
    elif isinstance(fn, dist.GaussianRandomWalk):
        lp = od.gaussian_random_walk_log_prob(v, p["scale"])
    elif isinstance(fn, dist.BernoulliLogits):
        lp = od.bernoulli_logits_log_prob(p["logits"], v)
    else:
        raise AssertionError(type(fn))
    return float(np.sum(lp))


def _interpreted_potential(model, args, sites, z):
    """U(z) by running the model on the constrained values of flat z."""
    vals, o, ladj = {}, 0, 0.0
    for name, size, tf in sites:
        v = z[o:o + size]
        if tf == "exp":
            ladj += float(np.sum(v))
            v = np.exp(v)
        vals[name] = torch.as_tensor(v if size > 1 else v[0], dtype=torch.float64)
        o += size
    with trace() as tr, substitute(data=vals):
        model(*args)
    return -(sum(_site_log_prob(s) for s in tr.values() if s["type"] == "sample") + ladj)


def _covtype_model(data, labels):
    # examples/covtype.py:66-71
    dim = data.shape[1]
    coefs = npa.sample("coefs", dist.Normal(torch.zeros(dim), torch.ones(dim)))
    with npa.plate("N", data.shape[0]) as idx:
        logits = torch.matmul(data[idx], coefs)
        return npa.sample("obs", dist.Bernoulli(logits=logits), obs=labels[idx])


def _logistic_data(N=300, D=5, seed=0):
    rng = np.random.default_rng(seed)
    X = torch.as_tensor(rng.normal(size=(N, D)))
    y = torch.as_tensor((rng.uniform(size=N) < 0.5).astype(np.float64))
    return X, y


def test_covtype_model_is_logistic():
    X, y = _logistic_data()
    kind, p, tr = match_model(_covtype_model, X, y)
    assert kind == "logistic"
    assert p["prior_scale"] == 1.0
    assert torch.equal(p["X"], X) and torch.equal(p["y"], y)
    assert p["sites"] == (("coefs", 5, None),)
    z = np.random.default_rng(1).normal(size=5)
    want = om.logistic_potential(z, X.numpy(), y.numpy(), 1.0)
    got = _interpreted_potential(_covtype_model, (X, y), p["sites"], z)
    np.testing.assert_allclose(got, want, rtol=1e-10)


def test_logistic_variants_and_deterministic():
    # test/infer/test_mcmc.py:104-168: sum(coefs * data, -1) and a deterministic site
    X, y = _logistic_data(200, 3, 2)

    def model(labels):
        coefs = npa.sample("coefs", dist.Normal(0.0, 2.5).expand([3]).to_event(1))
        logits = npa.deterministic("logits", torch.sum(coefs * X, dim=-1))
        return npa.sample("obs", dist.Bernoulli(logits=logits), obs=labels)

    kind, p, tr = match_model(model, y)
    assert kind == "logistic" and p["prior_scale"] == 2.5
    assert "logits" in tr and tr["logits"]["type"] == "deterministic"
    z = np.random.default_rng(3).normal(size=3)
    np.testing.assert_allclose(_interpreted_potential(model, (y,), p["sites"], z),
                               om.logistic_potential(z, X.numpy(), y.numpy(), 2.5), rtol=1e-10)


def _sv_model(returns):
    # examples/stochastic_volatility.py:58-66
    step_size = npa.sample("sigma", dist.Exponential(50.0))
    s = npa.sample("s", dist.GaussianRandomWalk(scale=step_size, num_steps=returns.shape[0]))
    nu = npa.sample("nu", dist.Exponential(0.1))
    return npa.sample("r", dist.StudentT(df=nu, loc=0.0, scale=torch.exp(s)), obs=returns)


def test_sv_model():
    r = torch.as_tensor(np.random.default_rng(4).standard_t(5, size=40) * 0.02)
    kind, p, _ = match_model(_sv_model, r)
    assert kind == "sv" and not p["noncentred"]
    assert (p["sigma_rate"], p["nu_rate"]) == (50.0, 0.1)
    assert p["sites"] == (("nu", 1, "exp"), ("s", 40, None), ("sigma", 1, "exp"))
    rng = np.random.default_rng(5)
    z = np.concatenate([[np.log(4.0)], np.cumsum(rng.normal(size=40) * 0.05) - 4.0, [-3.0]])
    np.testing.assert_allclose(_interpreted_potential(_sv_model, (r,), p["sites"], z),
                               om.sv_potential(z, r.numpy()), rtol=1e-10)


def test_sv_noncentred_other_rates_and_names():
    T = 30
    r = torch.as_tensor(np.random.default_rng(6).normal(size=T) * 0.03)

    def model(returns):
        vol = npa.sample("vol", dist.Exponential(20.0))
        dof = npa.sample("dof", dist.Exponential(0.5))
        inc = npa.sample("inc", dist.Normal(0.0, 1.0).expand([T]))
        s = npa.deterministic("s", vol * torch.cumsum(inc, dim=-1))
        npa.sample("r", dist.StudentT(dof, 0.0, torch.exp(s)), obs=returns)

    kind, p, tr = match_model(model, r)
    assert kind == "sv" and p["noncentred"]
    assert (p["sigma_rate"], p["nu_rate"]) == (20.0, 0.5)
    assert p["sites"] == (("dof", 1, "exp"), ("vol", 1, "exp"), ("inc", T, None))
    z = np.concatenate([[0.3, -3.5], np.random.default_rng(7).normal(size=T)])
    np.testing.assert_allclose(_interpreted_potential(model, (r,), p["sites"], z),
                               om.sv_nc_potential(z, r.numpy(), 20.0, 0.5), rtol=1e-10)


def test_hier_model():
    X, y = _logistic_data(150, 6, 8)

    def model(X, y):
        mu = npa.sample("mu", dist.Normal(0.0, 1.0))
        tau = npa.sample("tau", dist.HalfCauchy(1.0))
        eta = npa.sample("eta", dist.Normal(torch.zeros(6), 1.0))
        npa.sample("y", dist.Bernoulli(logits=X @ (mu + tau * eta)), obs=y)

    kind, p, _ = match_model(model, X, y)
    assert kind == "hier"
    assert p["sites"] == (("eta", 6, None), ("mu", 1, None), ("tau", 1, "exp"))
    z = np.concatenate([np.random.default_rng(9).normal(size=6), [0.2, -0.4]])
    np.testing.assert_allclose(_interpreted_potential(model, (X, y), p["sites"], z),
                               om.hier_potential(z, X.numpy(), y.numpy()), rtol=1e-10)


def test_interpreter_catches_a_planted_mistake():
    """The pin is not vacuous: the oracle potential with a wrong prior scale
    (a plausible matcher bug) disagrees with the interpreted model."""
    X, y = _logistic_data()
    _, p, _ = match_model(_covtype_model, X, y)
    z = np.random.default_rng(1).normal(size=5)
    got = _interpreted_potential(_covtype_model, (X, y), p["sites"], z)
    assert abs(got - om.logistic_potential(z, X.numpy(), y.numpy(), 1.5)) > 1e-3


@pytest.mark.parametrize("which", ["gaussian_lik", "prior_loc", "subsample", "probs", "two_x"])
def test_unsupported_models_are_refused(which):
    X, y = _logistic_data(50, 3)

    def model():
        if which == "prior_loc":
            b = npa.sample("b", dist.Normal(1.0, 1.0).expand([3]))
        else:
            b = npa.sample("b", dist.Normal(0.0, 1.0).expand([3]))
        if which == "gaussian_lik":
            npa.sample("y", dist.Normal(X @ b, 1.0), obs=y)
        elif which == "subsample":
            with npa.plate("N", 50, subsample_size=10):
                npa.sample("y", dist.Normal(X @ b, 1.0), obs=y)
        elif which == "probs":
            npa.sample("y", dist.Bernoulli(probs=X @ b), obs=y)
        elif which == "two_x":
            npa.sample("y", dist.Bernoulli(logits=X @ b + X @ b), obs=y)
        else:
            npa.sample("y", dist.Bernoulli(logits=X @ b), obs=y)

    with pytest.raises(NotImplementedError):
        match_model(model)


def test_plate_shapes():
    def model():
        with npa.plate("a", 4):
            with npa.plate("b", 3):
                x = npa.sample("x", dist.Normal(0.0, 1.0))
        return x

    from numpyro_amd.compile import trace_model
    tr = trace_model(model)
    assert tr["x"]["shape"] == (3, 4)
    assert isinstance(tr["x"]["value"], npa.compile.Sym)


def test_latent_without_handler_raises():
    with pytest.raises(RuntimeError, match="has no value"):
        npa.sample("x", dist.Normal(0.0, 1.0))


def test_invalid_prior_scale_is_an_error():
    X, y = _logistic_data(20, 2)

    def model():
        b = npa.sample("b", dist.Normal(0.0, -1.0).expand([2]))
        npa.sample("y", dist.Bernoulli(logits=X @ b), obs=y)

    with pytest.raises(ValueError, match="scale must be positive"):
        match_model(model)


def test_logistic_with_intercept():
    X, y = _logistic_data(120, 4, 12)

    def model(X, y):
        a = npa.sample("a", dist.Normal(0.0, 2.0))
        b = npa.sample("b", dist.Normal(0.0, 2.0).expand([4]).to_event(1))
        npa.sample("y", dist.Bernoulli(logits=b @ X.T + a), obs=y)

    # ravel_pytree flattens the latent dict in sorted site order: [a, b]
    kind, p, _ = match_model(model, X, y)
    assert kind == "logistic" and p["prior_scale"] == 2.0
    assert p["sites"] == (("a", 1, None), ("b", 4, None))
    assert p["X"].shape == (120, 5) and bool((p["X"][:, 0] == 1).all())
    z = np.random.default_rng(13).normal(size=5)
    np.testing.assert_allclose(_interpreted_potential(model, (X, y), p["sites"], z),
                               om.logistic_potential(z, p["X"].numpy(), y.numpy(), 2.0),
                               rtol=1e-10)

    def model2(X, y):  # an intercept whose name sorts last
        b = npa.sample("b", dist.Normal(0.0, 2.0).expand([4]).to_event(1))
        c = npa.sample("c", dist.Normal(0.0, 2.0))
        npa.sample("y", dist.Bernoulli(logits=X @ b + c), obs=y)

    kind, p, _ = match_model(model2, X, y)
    assert p["sites"] == (("b", 4, None), ("c", 1, None))
    assert bool((p["X"][:, 4] == 1).all())
    np.testing.assert_allclose(_interpreted_potential(model2, (X, y), p["sites"], z),
                               om.logistic_potential(z, p["X"].numpy(), y.numpy(), 2.0),
                               rtol=1e-10)


def test_intercept_with_other_scale_is_refused():
    X, y = _logistic_data(50, 3)

    def model():
        a = npa.sample("a", dist.Normal(0.0, 10.0))
        b = npa.sample("b", dist.Normal(0.0, 1.0).expand([3]))
        npa.sample("y", dist.Bernoulli(logits=X @ b + a), obs=y)

    with pytest.raises(NotImplementedError):
        match_model(model)


def _covtype_subsample_model(data, labels, subsample_size=None):
    # examples/covtype.py:66-71
    dim = data.shape[1]
    coefs = npa.sample("coefs", dist.Normal(torch.zeros(dim), torch.ones(dim)))
    with npa.plate("N", data.shape[0], subsample_size=subsample_size) as idx:
        logits = torch.matmul(data[idx], coefs)
        return npa.sample("obs", dist.Bernoulli(logits=logits), obs=labels[idx])


def test_subsample_plate_is_recorded_for_hmcecs():
    X, y = _logistic_data(200, 4)
    kind, p, _ = match_model(_covtype_subsample_model, X, y, 25)
    assert kind == "logistic" and p["subsample"] == ("N", 200, 25)
    assert torch.equal(p["X"], X)
    kind, p, _ = match_model(_covtype_subsample_model, X, y)
    assert "subsample" not in p

    def with_subsample_primitive(data, labels):
        b = npa.sample("b", dist.Normal(0.0, 2.0).expand([4]).to_event(1))
        with npa.plate("rows", data.shape[0], subsample_size=10):
            xb = npa.subsample(data, event_dim=1)
            yb = npa.subsample(labels, event_dim=0)
            npa.sample("y", dist.Bernoulli(logits=xb @ b), obs=yb)

    kind, p, _ = match_model(with_subsample_primitive, X, y)
    assert p["subsample"] == ("rows", 200, 10) and p["prior_scale"] == 2.0
    with pytest.raises(ValueError):
        npa.plate("N", 10, subsample_size=11)
