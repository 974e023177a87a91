"""NUTS(model) through the model compiler on the GPU.

A compiled model runs the same HIP kernels as the matching built-in model, so
with the same keys the chains must be bit-identical to the built-in model's;
the new stochastic-volatility entry point with other prior rates is checked
against the float64 oracle (U and dU/dz)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _covtype_model(data, labels):
    # examples/covtype.py:66-71
    import numpyro_amd as npa
    from numpyro_amd import distributions as dist
    dim = data.shape[1]
    coefs = npa.sample("coefs", dist.Normal(torch.zeros(dim), torch.ones(dim)))
    with npa.plate("N", data.shape[0]) as idx:
        logits = torch.matmul(data[idx], coefs)
        return npa.sample("obs", dist.Bernoulli(logits=logits), obs=labels[idx])


def _logistic_data(N=2000, D=4, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(N, D))
    beta = np.arange(1.0, D + 1.0) * 0.5
    y = (rng.uniform(size=N) < 1 / (1 + np.exp(-X @ beta))).astype(np.float32)
    return torch.as_tensor(X, dtype=torch.float32), torch.as_tensor(y)


def test_compiled_logistic_matches_builtin_bitwise():
    from numpyro_amd import random
    from numpyro_amd.infer import MCMC, NUTS
    from numpyro_amd.models import LogisticRegression
    X, y = _logistic_data()
    a = MCMC(NUTS(_covtype_model), num_warmup=200, num_samples=200, num_chains=4)
    a.run(random.PRNGKey(3), X, y)
    b = MCMC(NUTS(LogisticRegression(X, y)), num_warmup=200, num_samples=200, num_chains=4)
    b.run(random.PRNGKey(3))
    out = [m.get_samples(group_by_chain=True)["coefs"].cpu() for m in (a, b)]
    assert out[0].shape == (4, 200, 4)
    assert torch.equal(out[0], out[1])
    np.testing.assert_allclose(out[0].reshape(-1, 4).mean(0).numpy(), np.arange(1, 5) * 0.5,
                               atol=0.3)


def test_compiled_deterministic_and_init_params():
    import numpyro_amd as npa
    from numpyro_amd import distributions as dist
    from numpyro_amd import random
    from numpyro_amd.infer import MCMC, NUTS
    X, y = _logistic_data(500, 3, 1)
    Xc = X.cuda()

    def model(labels):
        coefs = npa.sample("coefs", dist.Normal(0.0, 1.0).expand([3]))
        logits = npa.deterministic("logits", torch.sum(coefs * Xc, dim=-1))
        return npa.sample("obs", dist.Bernoulli(logits=logits), obs=labels)

    mcmc = MCMC(NUTS(model), num_warmup=100, num_samples=50, num_chains=2)
    mcmc.run(random.PRNGKey(0), y, init_params={"coefs": torch.zeros(2, 3)})
    s = mcmc.get_samples()
    assert set(s) == {"coefs", "logits"}
    assert s["logits"].shape == (100, 500)
    torch.testing.assert_close(s["logits"], s["coefs"] @ Xc.T, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("noncentred", [False, True])
@pytest.mark.parametrize("rates", [(50.0, 0.1), (7.0, 0.35)])
def test_sv_rates_potential_vs_oracle(noncentred, rates):
    from numpyro_amd.models import StochasticVolatility, potential_energy
    from numpyro_amd.workloads import make_sv
    from oracle import models as om
    T, C = 300, 6
    r, _ = make_sv(T, 3)
    m = StochasticVolatility(r, noncentred=noncentred, sigma_rate=rates[0], nu_rate=rates[1])
    rng = np.random.default_rng(11)
    if noncentred:
        Z = np.concatenate([rng.normal(1.0, 0.5, (C, 1)), rng.normal(-3.5, 0.3, (C, 1)),
                            rng.normal(size=(C, T))], 1)
    else:
        Z = np.concatenate([rng.normal(1.0, 0.5, (C, 1)),
                            np.cumsum(rng.normal(0, 0.05, (C, T)), 1) - 4.0,
                            rng.normal(-3.5, 0.3, (C, 1))], 1)
    z = torch.tensor(Z, dtype=torch.float32, device="cuda", requires_grad=True)
    pe = potential_energy(m, z)
    pe.sum().backward()
    g = z.grad.cpu().numpy()
    Z32 = z.detach().cpu().numpy().astype(np.float64)
    f = om.sv_nc_potential_grad if noncentred else om.sv_potential_grad
    for c in range(C):
        u, gr = f(Z32[c], r.astype(np.float64), *rates)
        np.testing.assert_allclose(pe[c].item(), u, rtol=1e-6)
        np.testing.assert_allclose(g[c], gr, rtol=2e-4, atol=2e-4 * (np.abs(gr).max() + 1))


def test_compiled_sv_and_hier_match_builtin():
    import numpyro_amd as npa
    from numpyro_amd import distributions as dist
    from numpyro_amd import random
    from numpyro_amd.infer import MCMC, NUTS
    from numpyro_amd.models import HierarchicalLogistic, StochasticVolatility
    from numpyro_amd.workloads import make_sv

    def sv(returns):
        # examples/stochastic_volatility.py:58-66
        step_size = npa.sample("sigma", dist.Exponential(50.0))
        s = npa.sample("s", dist.GaussianRandomWalk(scale=step_size, num_steps=returns.shape[0]))
        nu = npa.sample("nu", dist.Exponential(0.1))
        return npa.sample("r", dist.StudentT(df=nu, loc=0.0, scale=torch.exp(s)), obs=returns)

    r, _ = make_sv(200, 3)
    rt = torch.as_tensor(r)
    a = MCMC(NUTS(sv), num_warmup=60, num_samples=30, num_chains=2)
    a.run(random.PRNGKey(5), rt)
    b = MCMC(NUTS(StochasticVolatility(r)), num_warmup=60, num_samples=30, num_chains=2)
    b.run(random.PRNGKey(5))
    sa, sb = a.get_samples(), b.get_samples()
    for k in ("nu", "s", "sigma"):
        assert torch.equal(sa[k], sb[k]), k

    X, y = _logistic_data(400, 5, 2)

    def hier(X, y):
        mu = npa.sample("mu", dist.Normal(0.0, 1.0))
        tau = npa.sample("tau", dist.HalfCauchy(1.0))
        eta = npa.sample("eta", dist.Normal(torch.zeros(5), 1.0))
        npa.sample("y", dist.Bernoulli(logits=X @ (mu + tau * eta)), obs=y)

    # the same given starting points (the built-in ragged model has no named
    # sites, so its default initial draws differ from the traced model's)
    z0 = torch.from_numpy(np.random.default_rng(1).uniform(-2, 2, (4, 7)).astype(np.float32))
    a = MCMC(NUTS(hier), num_warmup=60, num_samples=30, num_chains=4)
    a.run(random.PRNGKey(6), X, y, init_params={"eta": z0[:, :5], "mu": z0[:, 5],
                                                 "tau": z0[:, 6]})
    b = MCMC(NUTS(HierarchicalLogistic([X], [y], chains_per_group=4)), num_warmup=60,
             num_samples=30, num_chains=4)
    b.run(random.PRNGKey(6), init_params=z0)
    sa, sb = a.get_samples(), b.get_samples()
    assert torch.equal(sa["mu"], sb["mu"]) and torch.equal(sa["tau"], sb["tau"])
    assert torch.equal(sa["eta"], sb["eta"][:, :5])


def test_compiled_intercept_matches_augmented_builtin():
    import numpyro_amd as npa
    from numpyro_amd import distributions as dist
    from numpyro_amd import random
    from numpyro_amd.infer import MCMC, NUTS
    from numpyro_amd.models import LogisticRegression
    X, y = _logistic_data(1500, 3, 4)

    def model(X, y):
        b = npa.sample("b", dist.Normal(0.0, 1.5).expand([3]).to_event(1))
        a = npa.sample("a", dist.Normal(0.0, 1.5))
        npa.sample("y", dist.Bernoulli(logits=X @ b + a), obs=y)

    # the flat latent is [a, b] (sorted site names, as ravel_pytree), so a
    # flat inverse_mass_matrix applies its first entry to the intercept; both
    # start from the same given point (model-based init draws site by site in
    # the model's order, the built-in model has one site)
    imm = torch.tensor([0.5, 1.0, 2.0, 3.0])
    z0 = torch.tensor([[0.1, -0.2, 0.3, 0.0], [-0.3, 0.2, 0.1, -0.1]])
    m = MCMC(NUTS(model, inverse_mass_matrix=imm, adapt_mass_matrix=False), num_warmup=100,
             num_samples=100, num_chains=2)
    m.run(random.PRNGKey(8), X, y, init_params={"a": z0[:, 0], "b": z0[:, 1:]})
    Xa = torch.cat([torch.ones(X.shape[0], 1), X], 1)
    r = MCMC(NUTS(LogisticRegression(Xa, y, prior_scale=1.5), inverse_mass_matrix=imm,
                  adapt_mass_matrix=False), num_warmup=100, num_samples=100, num_chains=2)
    r.run(random.PRNGKey(8), init_params=z0)
    s, t = m.get_samples(), r.get_samples()["coefs"]
    assert set(s) == {"a", "b"}
    assert torch.equal(s["a"], t[:, 0]) and torch.equal(s["b"], t[:, 1:])
