"""HMCECS on the GPU against the oracle (oracle/ecs.py): the subsampled
potential and gradient, the key-split / randint draws, one Gibbs update of
every chain (keys, indices, acceptance, U and dU bit-for-bit or to fp32
tolerance), and sampling of the covtype model with a subsample plate."""
import numpy as np
import pytest
import torch

from oracle import ecs
from oracle.jax_random import split

pytestmark = pytest.mark.gpu


def _data(N, D, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(N, D)).astype(np.float32)
    beta = np.linspace(-1.0, 1.0, D)
    y = (rng.uniform(size=N) < 1 / (1 + np.exp(-X.astype(np.float64) @ beta))).astype(np.float32)
    return X, y, beta


def _model(X, y, m, degree, ref, C, seed=0, nb=1):
    from numpyro_amd.models import SubsampledLogisticRegression
    mdl = SubsampledLogisticRegression(X, y, m, prior_scale=1.3, num_blocks=nb,
                                       proxy_degree=degree, reference=ref)
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, X.shape[0], (C, m)).astype(np.int32)
    keys = rng.integers(0, 2 ** 32, (C, 2), dtype=np.uint64).astype(np.uint32)
    mdl.idx = torch.as_tensor(idx).cuda()
    mdl.keys = torch.as_tensor(keys.view(np.int32)).cuda()
    return mdl, idx, keys


@pytest.mark.parametrize("degree", [0, 1, 2])
@pytest.mark.parametrize("N,D,m", [(3000, 8, 100), (5000, 33, 1000), (20000, 256, 700),
                                   (4000, 130, 300), (3000, 300, 200)])
def test_subsampled_potential_vs_oracle(degree, N, D, m):
    from numpyro_amd.models import potential_energy
    X, y, beta = _data(N, D)
    ref = beta + 0.02
    C = 5
    mdl, idx, _ = _model(X, y, m, degree, ref if degree else None, C)
    st = ecs.taylor_stats(X, y, ref, degree) if degree else None
    Z = beta + np.random.default_rng(3).normal(0, 0.05, (C, D))
    z = torch.tensor(Z, dtype=torch.float32, device="cuda", requires_grad=True)
    pe = potential_energy(mdl, z)
    pe.sum().backward()
    g = z.grad.cpu().numpy()
    Z32 = z.detach().cpu().numpy().astype(np.float64)
    for c in range(C):
        u, gr = ecs.potential_grad(Z32[c], idx[c], X, y, N, st, 1.3)
        np.testing.assert_allclose(pe[c].item(), u, rtol=2e-6)
        np.testing.assert_allclose(g[c], gr, rtol=1e-4, atol=1e-4 * np.abs(gr).max())


def test_randint_and_split_keys_match_oracle():
    from numpyro_amd import _lib
    from numpyro_amd import random as nrandom
    part = nrandom.PARTITIONABLE
    rng = np.random.default_rng(0)
    keys = rng.integers(0, 2 ** 32, (7, 2), dtype=np.uint64).astype(np.uint32)
    kd = torch.as_tensor(keys.view(np.int32)).cuda()
    lib = _lib.load()
    for n, lo, hi in [(1, 0, 5), (37, 0, 2_000_000), (300, -4, 70000), (5, 3, 3)]:
        out = torch.empty((7, n), dtype=torch.int32, device="cuda")
        _lib.check(lib.nax_randint(_lib.ptr(kd), 7, n, lo, hi, int(part), _lib.ptr(out)), "randint")
        for c in range(7):
            np.testing.assert_array_equal(out[c].cpu().numpy(), ecs.randint(keys[c], n, lo, hi, part))
    out = torch.empty((7, 3, 2), dtype=torch.int32, device="cuda")
    _lib.check(lib.nax_split_keys(_lib.ptr(kd), 7, 3, int(part), _lib.ptr(out)), "split")
    got = out.cpu().numpy().view(np.uint32)
    for c in range(7):
        np.testing.assert_array_equal(got[c], split(keys[c], 3, part))


@pytest.mark.parametrize("degree,nb", [(2, 4), (0, 1), (1, 7)])
def test_gibbs_step_vs_oracle(degree, nb):
    from numpyro_amd import _lib
    from numpyro_amd import random as nrandom
    from numpyro_amd.infer.hmc import Engine
    N, D, m, C = 4000, 6, 60, 24
    X, y, beta = _data(N, D, 2)
    ref = beta + 0.05
    mdl, idx, keys = _model(X, y, m, degree, ref if degree else None, C, seed=4, nb=nb)
    st = ecs.taylor_stats(X, y, ref, degree) if degree else None
    eng = Engine(mdl, C, "NUTS", 10)
    Z = beta + np.random.default_rng(5).normal(0, 0.1, (C, D))
    z0 = torch.tensor(Z, dtype=torch.float32)
    init_keys = np.random.default_rng(6).integers(0, 2 ** 32, (C, 2), dtype=np.uint64)
    eng.init(z0, init_keys.astype(np.uint32))
    s0 = eng.get_state()
    pe0 = np.asarray(s0.potential_energy, dtype=np.float64)
    g0 = s0.z_grad.cpu().numpy()
    acc = torch.zeros(C, dtype=torch.float32, device="cuda")
    _lib.check(eng.lib.nax_ecs_gibbs(eng.h, _lib.ptr(acc)), "gibbs")
    s1 = eng.get_state()
    idx1 = mdl.idx.cpu().numpy()
    keys1 = mdl.keys.cpu().numpy().view(np.uint32)
    pe1 = np.asarray(s1.potential_energy, dtype=np.float64)
    g1 = s1.z_grad.cpu().numpy()
    zc = z0.numpy().astype(np.float64)
    part = nrandom.PARTITIONABLE
    n_acc = n_rej = 0
    for c in range(C):
        k, idx_o, pe_o, g_o, ap = ecs.gibbs_step(keys[c], idx[c], zc[c], pe0[c], X, y, N, nb, st,
                                                 1.3, part)
        np.testing.assert_array_equal(keys1[c], k)
        np.testing.assert_allclose(acc[c].item(), ap, rtol=1e-3, atol=1e-6)
        u = float(ecs.uniform(k, None, 0.0, 1.0, part))
        if abs(u - ap) < 1e-3:
            continue  # fp32 vs fp64 potentials may decide either way
        np.testing.assert_array_equal(idx1[c], idx_o)
        if g_o is None:
            n_rej += 1
            assert pe1[c] == pe0[c]
            np.testing.assert_array_equal(g1[c], g0[c])
        else:
            n_acc += 1
            np.testing.assert_allclose(pe1[c], pe_o, rtol=2e-6)
            np.testing.assert_allclose(g1[c], g_o, rtol=1e-4, atol=1e-4 * np.abs(g_o).max())
    print(f"degree {degree} blocks {nb}: {n_acc} accepted, {n_rej} rejected")
    assert n_acc + n_rej >= C - 3


def test_gibbs_step_rejects_nan_proposals():
    # hmc_gibbs.py:660-661: accept_prob = clip(exp(pe - pe_new), None, 1) keeps
    # a NaN, and bernoulli(key, NaN) rejects.  Rows 0..9 of X hold a NaN; the
    # chains start on subsamples without them, so a proposal that draws one
    # has a NaN potential and must be rejected (the accept_prob stays NaN).
    from numpyro_amd import _lib
    from numpyro_amd import random as nrandom
    from numpyro_amd.infer.hmc import Engine
    N, D, m, C = 400, 4, 20, 32
    X, y, beta = _data(N, D, 8)
    X[:10, 1] = np.nan
    mdl, idx, keys = _model(X, y, m, 0, None, C, seed=9, nb=1)
    idx = np.random.default_rng(10).integers(10, N, (C, m)).astype(np.int32)
    mdl.idx = torch.as_tensor(idx).cuda()
    eng = Engine(mdl, C, "NUTS", 10)
    zc = beta + np.random.default_rng(11).normal(0, 0.1, (C, D))
    eng.init(torch.tensor(zc, dtype=torch.float32),
             np.random.default_rng(12).integers(0, 2 ** 32, (C, 2), dtype=np.uint64)
             .astype(np.uint32))
    pe0 = np.asarray(eng.get_state().potential_energy, dtype=np.float64)
    assert np.isfinite(pe0).all()
    acc = torch.zeros(C, dtype=torch.float32, device="cuda")
    _lib.check(eng.lib.nax_ecs_gibbs(eng.h, _lib.ptr(acc)), "gibbs")
    idx1 = mdl.idx.cpu().numpy()
    pe1 = np.asarray(eng.get_state().potential_energy, dtype=np.float64)
    acc = acc.cpu().numpy()
    z32 = torch.tensor(zc, dtype=torch.float32).numpy().astype(np.float64)
    n_nan = 0
    for c in range(C):
        k, idx_o, pe_o, g_o, ap = ecs.gibbs_step(keys[c], idx[c], z32[c], pe0[c], X, y, N, 1,
                                                 None, 1.3, nrandom.PARTITIONABLE)
        if np.isnan(ap):
            n_nan += 1
            assert np.isnan(acc[c]) and g_o is None
            np.testing.assert_array_equal(idx1[c], idx[c])
            assert pe1[c] == pe0[c]
    assert n_nan >= 3, n_nan


def _covtype(data, labels, subsample_size=None):
    # examples/covtype.py:66-71
    import numpyro_amd as npa
    from numpyro_amd import distributions as dist
    dim = data.shape[1]
    coefs = npa.sample("coefs", dist.Normal(torch.zeros(dim), torch.ones(dim)))
    with npa.plate("N", data.shape[0], subsample_size=subsample_size) as idx:
        logits = torch.matmul(data[idx], coefs)
        return npa.sample("obs", dist.Bernoulli(logits=logits), obs=labels[idx])


@pytest.mark.parametrize("proxy", [True, False])
def test_hmcecs_covtype_posterior(proxy):
    # examples/covtype.py:154-165 on a synthetic design; the posterior mean
    # must agree with full-data NUTS (as test/infer/test_hmc_gibbs.py:302-330
    # checks HMCECS against the true location).
    from numpyro_amd import random
    from numpyro_amd.infer import HMCECS, MCMC, NUTS, init_to_value
    X, y, beta = _data(20000, 5, 7)
    Xt, yt = torch.as_tensor(X), torch.as_tensor(y)
    full = MCMC(NUTS(_covtype), num_warmup=300, num_samples=300, num_chains=4)
    full.run(random.PRNGKey(1), Xt, yt)
    ref_mean = full.get_samples()["coefs"].mean(0).cpu().numpy()
    prox = HMCECS.taylor_proxy({"coefs": torch.as_tensor(ref_mean)}) if proxy else None
    # examples/covtype.py:154-165: the inner kernel starts at the reference point
    init = init_to_value(values={"coefs": torch.as_tensor(ref_mean)}) if proxy else None
    kern = HMCECS(NUTS(_covtype, init_strategy=init), num_blocks=10, proxy=prox)
    mcmc = MCMC(kern, num_warmup=300, num_samples=300, num_chains=4)
    mcmc.run(random.PRNGKey(2), Xt, yt, 1000)
    s = mcmc.get_samples()["coefs"].cpu().numpy()
    assert s.shape == (1200, 5) and np.isfinite(s).all()
    sd = full.get_samples()["coefs"].std(0).cpu().numpy()
    if proxy:
        assert np.all(np.abs(s.mean(0) - ref_mean) < 4 * sd), (s.mean(0), ref_mean, sd)
    else:
        # Without a control variate the n/m-scaled log-likelihood estimate has
        # a standard deviation of hundreds of nats here, and the chains favour
        # subsamples that overstate it (the bias of naive subsampling the
        # reference's docstring warns of, hmc_gibbs.py:509-511): only the
        # direction of the coefficients is checked.
        assert np.all(np.sign(s.mean(0)[[0, 1, 3, 4]]) == np.sign(ref_mean[[0, 1, 3, 4]]))
        assert np.all(np.abs(s.mean(0) - ref_mean) < 1.0), (s.mean(0), ref_mean)
    acc = kern.gibbs_accept_prob.cpu().numpy()
    assert acc.shape == (4,) and np.all((acc >= 0) & (acc <= 1))
    idx = kern.subsample_indices["N"]
    assert idx.shape == (4, 1000) and int(idx.min()) >= 0 and int(idx.max()) < 20000
    # the Gibbs site is collected with the samples (z = {**z_gibbs, **hmc_state.z})
    drawn = mcmc.get_samples(group_by_chain=True)["N"]
    assert drawn.shape == (4, 300, 1000) and drawn.dtype == torch.int32
    assert torch.equal(drawn[:, -1], idx)
    assert bool((drawn[:, 1:] != drawn[:, :-1]).any())
    mcmc.print_summary()  # the integer Gibbs site is left out of the summary


def test_hmcecs_with_hmc_and_first_order_proxy():
    # test/infer/test_hmc_gibbs.py:302-330 runs HMCECS over HMC and NUTS with
    # Taylor proxies of degree 1 and 2
    from numpyro_amd import random
    from numpyro_amd.infer import HMC, HMCECS, MCMC, NUTS
    X, y, beta = _data(20000, 4, 9)
    Xt, yt = torch.as_tensor(X), torch.as_tensor(y)
    full = MCMC(NUTS(_covtype), num_warmup=200, num_samples=200, num_chains=2)
    full.run(random.PRNGKey(3), Xt, yt)
    ref = full.get_samples()["coefs"]
    kern = HMCECS(HMC(_covtype, trajectory_length=1.0), num_blocks=20,
                  proxy=HMCECS.taylor_proxy({"coefs": ref.mean(0)}, degree=1))
    mcmc = MCMC(kern, num_warmup=200, num_samples=200, num_chains=2)
    mcmc.run(random.PRNGKey(4), Xt, yt, 2000)
    s = mcmc.get_samples()["coefs"]
    sd = ref.std(0)
    assert bool(((s.mean(0) - ref.mean(0)).abs() < 4 * sd).all()), (s.mean(0), ref.mean(0), sd)


def test_plain_nuts_refuses_a_subsample_plate():
    from numpyro_amd import random
    from numpyro_amd.infer import MCMC, NUTS
    X, y, _ = _data(500, 3)
    mcmc = MCMC(NUTS(_covtype), num_warmup=5, num_samples=5)
    with pytest.raises(ValueError, match="HMCECS"):
        mcmc.run(random.PRNGKey(0), torch.as_tensor(X), torch.as_tensor(y), 50)


def test_init_to_value_starts_chains_at_the_value():
    from numpyro_amd import random
    from numpyro_amd.infer import MCMC, NUTS, init_to_value
    X, y, _ = _data(2000, 3, 11)
    start = torch.tensor([0.3, -0.2, 0.1])
    kern = NUTS(_covtype, init_strategy=init_to_value(values={"coefs": start}))
    keys = np.stack([np.asarray(random.PRNGKey(i), np.uint32) for i in range(3)])
    kern.init(keys, 10, None, (torch.as_tensor(X), torch.as_tensor(y)), {})
    z = kern._engine.get_state().z.cpu()
    assert torch.equal(z, start.expand(3, 3))
    mcmc = MCMC(kern, num_warmup=50, num_samples=50, num_chains=3)
    mcmc.run(random.PRNGKey(1), torch.as_tensor(X), torch.as_tensor(y))
    assert mcmc.get_samples()["coefs"].shape == (150, 3)


def test_estimated_potential_tracks_the_full_potential():
    # test/infer/test_hmc_gibbs.py:402-441 with the logistic likelihood: at the
    # drawn points, exp(U_full - U_estimated) stays near 1
    from numpyro_amd import random
    from numpyro_amd.infer import HMCECS, MCMC, NUTS
    from numpyro_amd.models import LogisticRegression, potential_energy
    X, y, beta = _data(2000, 3, 13)
    Xt, yt = torch.as_tensor(X), torch.as_tensor(y)
    kern = HMCECS(NUTS(_covtype), num_blocks=20,
                  proxy=HMCECS.taylor_proxy({"coefs": torch.as_tensor(beta, dtype=torch.float32)}))
    mcmc = MCMC(kern, num_warmup=200, num_samples=200, num_chains=2)
    mcmc.run(random.PRNGKey(0), Xt, yt, 100, extra_fields=("hmc_state.potential_energy",))
    pes = mcmc.get_extra_fields()["hmc_state.potential_energy"].double()
    z = mcmc.get_samples()["coefs"]
    full = potential_energy(LogisticRegression(Xt, yt), z).double()
    r = torch.exp(full - pes)
    assert float(r.var()) < 1.0, (float(r.mean()), float(r.var()))


def test_hmcecs_extra_fields_accept_prob_is_the_gibbs_acceptance():
    # HMCECSState (hmc_gibbs.py:490-491, 681): accept_prob is the Gibbs
    # block acceptance of each iteration, hmc_state.accept_prob the inner
    # kernel's; a bare HMCState field does not exist on HMCECSState
    from numpyro_amd import random
    from numpyro_amd.infer import HMCECS, MCMC, NUTS
    X, y, beta = _data(3000, 3, 21)
    Xt, yt = torch.as_tensor(X), torch.as_tensor(y)
    prox = HMCECS.taylor_proxy({"coefs": torch.as_tensor(beta, dtype=torch.float32)})
    kern = HMCECS(NUTS(_covtype), num_blocks=10, proxy=prox)
    mcmc = MCMC(kern, num_warmup=50, num_samples=40, num_chains=3)
    mcmc.run(random.PRNGKey(5), Xt, yt, 200,
             extra_fields=("accept_prob", "hmc_state.accept_prob"))
    ex = mcmc.get_extra_fields(group_by_chain=True)
    ga, ha = ex["accept_prob"], ex["hmc_state.accept_prob"]
    assert ga.shape == (3, 40) and ha.shape == (3, 40)
    assert bool(((ga >= 0) & (ga <= 1)).all())
    assert torch.equal(ga[:, -1].cpu(), kern.gibbs_accept_prob.cpu())
    assert not torch.equal(ga.cpu(), ha.cpu())
    mcmc2 = MCMC(HMCECS(NUTS(_covtype), num_blocks=10, proxy=prox), num_warmup=5,
                 num_samples=5)
    with pytest.raises(AttributeError, match="HMCECSState"):
        mcmc2.run(random.PRNGKey(6), Xt, yt, 200, extra_fields=("diverging",))
