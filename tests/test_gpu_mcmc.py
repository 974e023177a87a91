"""Statistical and API tests of the MCMC driver on the GPU, re-expressing
test/infer/test_mcmc.py (unnormalised normal 28-72, logistic regression
104-168, divergences 488-509, extra fields 688-700, functional API 703-727)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_unnormalized_normal_potential_fn():
    # test_mcmc.py:28-72 with a torch potential_fn
    from numpyro_amd import random
    from numpyro_amd.infer import MCMC, NUTS
    true_mean, true_std = 1.0, 0.5

    def potential_fn(z):
        return 0.5 * torch.sum(((z - true_mean) / true_std) ** 2)

    kernel = NUTS(potential_fn=potential_fn)
    mcmc = MCMC(kernel, num_warmup=1000, num_samples=4000, num_chains=4, progress_bar=False)
    mcmc.run(random.PRNGKey(0), init_params=torch.zeros(4, 1))
    s = mcmc.get_samples().cpu().numpy()
    np.testing.assert_allclose(s.mean(), true_mean, rtol=0.07)
    np.testing.assert_allclose(s.std(), true_std, rtol=0.07)


@pytest.mark.parametrize("algo", ["NUTS", "HMC"])
def test_logistic_regression(algo):
    # test_mcmc.py:104-168: N=3000, dim=3, coefs [1,2,3], atol 0.4
    from numpyro_amd import random
    from numpyro_amd.infer import HMC, MCMC, NUTS
    from numpyro_amd.models import LogisticRegression
    rng = np.random.default_rng(0)
    N, dim = 3000, 3
    data = rng.normal(size=(N, dim))
    true_coefs = np.arange(1.0, dim + 1.0)
    labels = (rng.uniform(size=N) < 1 / (1 + np.exp(-data @ true_coefs))).astype(np.float32)
    m = LogisticRegression(data, labels)
    kernel = NUTS(m) if algo == "NUTS" else HMC(m, trajectory_length=8)
    mcmc = MCMC(kernel, num_warmup=1000, num_samples=2000, num_chains=2)
    mcmc.run(random.PRNGKey(2))
    coefs = mcmc.get_samples()["coefs"].cpu().numpy()
    assert coefs.shape == (4000, dim)
    np.testing.assert_allclose(coefs.mean(0), true_coefs, atol=0.4)


@pytest.mark.parametrize("adapt_step_size", [True, False])
def test_diverging(adapt_step_size):
    # test_mcmc.py:488-509 (loc ~ N(0,1), obs ~ N(loc,1), step_size=10)
    from numpyro_amd import random
    from numpyro_amd.infer import MCMC, NUTS
    data = torch.from_numpy(np.random.default_rng(0).normal(size=1000)).cuda()

    def potential_fn(z):
        return 0.5 * z[0] ** 2 + 0.5 * torch.sum((data - z[0]) ** 2)

    kernel = NUTS(potential_fn=potential_fn, step_size=10.0, adapt_step_size=adapt_step_size,
                  adapt_mass_matrix=False)
    num_warmup = num_samples = 300
    mcmc = MCMC(kernel, num_warmup=num_warmup, num_samples=num_samples)
    mcmc.warmup(random.PRNGKey(1), extra_fields=["diverging"], collect_warmup=True,
                init_params=torch.zeros(1))
    wd = int(mcmc.get_extra_fields()["diverging"].sum())
    mcmc.run(random.PRNGKey(2), extra_fields=["diverging"])
    nd = wd + int(mcmc.get_extra_fields()["diverging"].sum())
    if adapt_step_size:
        assert nd <= num_warmup
    else:
        assert nd == num_warmup + num_samples


def test_extra_fields_shapes():
    from numpyro_amd import random
    from numpyro_amd.infer import MCMC, NUTS
    kernel = NUTS(potential_fn=lambda z: 0.5 * torch.sum(z ** 2))
    mcmc = MCMC(kernel, num_warmup=200, num_samples=300)
    mcmc.run(random.PRNGKey(0), init_params=torch.zeros(5),
             extra_fields=("num_steps", "adapt_state.step_size", "accept_prob",
                           "potential_energy"))
    s = mcmc.get_samples(group_by_chain=True)
    assert s.shape == (1, 300, 5)
    st = mcmc.get_extra_fields(group_by_chain=True)
    assert st["num_steps"].shape == (1, 300)
    assert st["adapt_state.step_size"].shape == (1, 300)
    assert (st["num_steps"] >= 1).all()
    # mcmc.py:799-804: the same samples, now in host memory
    before = mcmc.get_samples().cpu()
    mcmc.transfer_states_to_host()
    after = mcmc.get_samples()
    assert after.device.type == "cpu" and torch.equal(before, after)
    assert mcmc.get_extra_fields()["num_steps"].device.type == "cpu"


def test_post_warmup_state_resume():
    # mcmc.py:558-587: run() after warmup() skips warmup
    from numpyro_amd import random
    from numpyro_amd.infer import MCMC, NUTS
    kernel = NUTS(potential_fn=lambda z: 0.5 * torch.sum((z - 3.0) ** 2))
    mcmc = MCMC(kernel, num_warmup=300, num_samples=500, num_chains=2)
    mcmc.warmup(random.PRNGKey(0), init_params=torch.zeros(2, 2))
    st = mcmc.post_warmup_state
    assert int(st.i.min()) == 300
    mcmc.run(random.PRNGKey(1))
    s = mcmc.get_samples().cpu().numpy()
    assert s.shape == (1000, 2)
    np.testing.assert_allclose(s.mean(0), [3.0, 3.0], atol=0.2)
    assert int(mcmc.last_state.i.min()) == 800


def test_functional_api_normal():
    # test_mcmc.py:703-774 (functional hmc + fori_collect)
    from numpyro_amd import random
    from numpyro_amd.infer import fori_collect, hmc
    true_mean, true_std = 1.0, 2.0
    init_kernel, sample_kernel = hmc(
        lambda z: 0.5 * torch.sum(((z - true_mean) / true_std) ** 2), algo="NUTS")
    keys = random.split(random.PRNGKey(0), 2)
    st = init_kernel(torch.tensor([[0.0], [-1.0]]), num_warmup=1000, rng_key=keys)
    s = fori_collect(1000, 4000, sample_kernel, st, transform=lambda x: x.z)
    s = s.cpu().numpy()
    assert s.shape == (2, 3000, 1)
    np.testing.assert_allclose(s.mean(axis=(1, 2)), [true_mean] * 2, rtol=0.15)
    np.testing.assert_allclose(s.std(axis=(1, 2)), [true_std] * 2, rtol=0.1)


def test_graph_replay_matches_direct_launches():
    # nax_run replays its tick batches as a captured hipGraph (after two
    # direct batches); with direct launches only, every draw must be the same
    # bit for bit.  ~60 iterations x ~7 leaves = hundreds of ticks: many
    # replays, each followed by the loop's own host synchronisation.
    import numpy as np
    import torch
    from numpyro_amd import random as nrandom
    from numpyro_amd.infer.hmc import Engine
    from numpyro_amd.models import LogisticRegression
    rng = np.random.default_rng(5)
    N, D, C = 500, 8, 64
    X = rng.normal(size=(N, D)).astype(np.float32)
    y = (rng.uniform(size=N) < 0.5).astype(np.float32)
    keys = nrandom.chain_init_keys(nrandom.PRNGKey(3), C)
    outs = []
    for graphs in (True, False):
        eng = Engine(LogisticRegression(X, y, kernel="mfma"), C, "NUTS", 30)
        eng.set_graphs(graphs)
        eng.init(torch.zeros((C, D)), keys)
        out = eng.run(60, collect_fields=("z", "num_steps", "potential_energy"), size=60)
        outs.append({k: v.cpu().numpy() for k, v in out.items()})
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k
    assert outs[0]["num_steps"].sum() > 16 * 3 * C  # enough ticks for several replays


def test_out_of_support_obs_cannot_initialise():
    # BernoulliLogits with labels outside {0, 1}: numpyro warns and
    # find_valid_initial_params fails (infer/util.py:366-508, 830-832)
    from numpyro_amd import random as nrandom
    from numpyro_amd.infer import MCMC, NUTS
    from numpyro_amd.models import LogisticRegression
    X = np.random.default_rng(0).normal(size=(50, 3)).astype(np.float32)
    y = np.full(50, 0.5, np.float32)
    with pytest.warns(UserWarning, match="Out-of-support"):
        m = LogisticRegression(X, y)
    mcmc = MCMC(NUTS(m), num_warmup=5, num_samples=5)
    with pytest.raises(RuntimeError, match="Cannot find valid initial parameters"):
        mcmc.run(nrandom.PRNGKey(0))
