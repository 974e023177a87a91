"""Stochastic-volatility potential + hand adjoint on the GPU against the
float64 oracle (whose gradients are pinned by central differences), whole
NUTS chains against the oracle, and centred vs non-centred posteriors."""
import numpy as np
import pytest
import torch

from oracle import hmc as ohmc
from oracle import jax_random as ojr
from oracle import models as om

pytestmark = pytest.mark.gpu


def _returns(T, seed):
    from numpyro_amd.workloads import make_sv
    r, _ = make_sv(T, seed)
    return r


def _points(T, C, noncentred, seed):
    rng = np.random.default_rng(seed)
    Z = np.zeros((C, T + 2), np.float32)
    for c in range(C):
        u_nu = np.log(rng.uniform(2, 20))
        u_sig = np.log(rng.uniform(0.01, 0.1))
        if noncentred:
            Z[c] = np.concatenate([[u_nu, u_sig], rng.normal(size=T)])
        else:
            s = np.cumsum(np.exp(u_sig) * rng.normal(size=T)) - 4.0
            Z[c] = np.concatenate([[u_nu], s, [u_sig]])
    return Z


@pytest.mark.parametrize("noncentred", [False, True])
@pytest.mark.parametrize("T, C, waves", [(3000, 16, None), (5, 3, None), (64, 4, None),
                                         (65, 5, None), (130, 8, None), (4097, 2, None),
                                         (3000, 16, 4), (3000, 16, 16), (5, 3, 2),
                                         (1025, 5, 16), (4097, 2, 8)])
def test_sv_potential_grad(noncentred, T, C, waves):
    # waves per chain fixed by the launch shape, else one
    from numpyro_amd import _lib
    from numpyro_amd.infer.hmc import Engine
    from numpyro_amd.models import StochasticVolatility
    r = _returns(T, 30 + T)
    Z = _points(T, C, noncentred, T + C)
    eng = Engine(StochasticVolatility(r, noncentred=noncentred), C, "NUTS", 0)
    if waves:
        eng.set_launch_shape(waves)
    z = torch.from_numpy(Z).cuda()
    pe = torch.empty(C, dtype=torch.float64, device="cuda")
    g = torch.empty((C, T + 2), dtype=torch.float32, device="cuda")
    _lib.check(_lib.load().nax_potential_grad(eng.h, _lib.ptr(z), T + 2, C, _lib.ptr(pe),
                                              _lib.ptr(g)))
    f = om.sv_nc_potential_grad if noncentred else om.sv_potential_grad
    for c in range(C):
        u, gr = f(Z[c].astype(np.float64), r.astype(np.float64))
        np.testing.assert_allclose(pe[c].item(), u, rtol=2e-6, atol=1e-3)
        np.testing.assert_allclose(g[c].cpu().numpy(), gr, rtol=2e-4,
                                   atol=2e-5 * (np.abs(gr).max() + 1))


@pytest.mark.parametrize("noncentred", [False, True])
def test_sv_nuts_chain_matches_oracle(noncentred):
    from numpyro_amd.infer import hmc
    from numpyro_amd.infer.util import fori_collect
    from numpyro_amd.models import StochasticVolatility
    T, C = 20, 2
    r = _returns(T, 7)
    Z = _points(T, C, noncentred, 8)
    keys = ojr.split(ojr.PRNGKey(3), C)
    kw = dict(step_size=0.05, adapt_step_size=False, adapt_mass_matrix=False)
    ik, sk = hmc(StochasticVolatility(r, noncentred=noncentred))
    st = ik(torch.from_numpy(Z), 0, rng_key=keys, **kw)
    gns, gz = fori_collect(0, 25, sk, st, transform=lambda s: (s.num_steps, s.z))
    gns, gz = gns.cpu().numpy(), gz.cpu().numpy()
    f = om.sv_nc_potential_grad if noncentred else om.sv_potential_grad
    for c in range(C):
        oik, osk = ohmc.hmc(lambda q: f(q, r.astype(np.float64)))
        ost = oik(Z[c].astype(np.float64), 0, rng_key=keys[c], **kw)
        ns, zs = [], []
        for _ in range(25):
            ost = osk(ost)
            ns.append(ost.num_steps)
            zs.append(ost.z)
        assert list(gns[c]) == ns
        np.testing.assert_allclose(gz[c], np.stack(zs), rtol=1e-3, atol=1e-3)


def test_sv_centred_vs_noncentred_posterior():
    from numpyro_amd import random
    from numpyro_amd.diagnostics import effective_sample_size
    from numpyro_amd.infer import MCMC, NUTS
    from numpyro_amd.models import StochasticVolatility
    T, C = 100, 64
    r = _returns(T, 11)
    res = {}
    for nc in (False, True):
        m = StochasticVolatility(r, noncentred=nc)
        mcmc = MCMC(NUTS(m), num_warmup=400, num_samples=400, num_chains=C)
        z0 = torch.zeros(C, T + 2)
        z0[:, 0] = np.log(5.0)
        z0[:, 1 if nc else T + 1] = np.log(0.05)
        mcmc.run(random.PRNGKey(1), init_params=z0)
        smp = mcmc.get_samples(group_by_chain=True)
        sig = smp["sigma"].cpu().numpy().astype(np.float64)
        s_last = smp["s"][..., -1].cpu().numpy().astype(np.float64)
        res[nc] = [(x.mean(), x.std(), effective_sample_size(x)) for x in (sig, s_last)]
    for (m0, s0, e0), (m1, s1, e1) in zip(res[False], res[True]):
        mcse = np.sqrt(s0 ** 2 / e0 + s1 ** 2 / e1)
        assert abs(m0 - m1) < 4 * mcse, (m0, m1, mcse)


@pytest.mark.parametrize("noncentred", [False, True])
@pytest.mark.parametrize("T, C", [(300, 6), (65, 3)])
def test_sv_softplus_potential_grad(noncentred, T, C):
    # sigma and nu unconstrained by SoftplusTransform (transforms.py:1156-1175):
    # U against the float64 oracle, dU against central differences of it
    from numpyro_amd import _lib
    from numpyro_amd.infer.hmc import Engine
    from numpyro_amd.models import StochasticVolatility
    r = _returns(T, 70 + T)
    Z = _points(T, C, noncentred, 3 * T + C)
    pos = [0, 1] if noncentred else [0, T + 1]
    # unconstrained coordinates: softplus^-1 of the same sigma, nu values
    x = np.exp(Z[:, pos].astype(np.float64))
    Z[:, pos] = (np.log(-np.expm1(-x)) + x).astype(np.float32)
    eng = Engine(StochasticVolatility(r, noncentred=noncentred, positive_transform="softplus"),
                 C, "NUTS", 0)
    z = torch.from_numpy(Z).cuda()
    pe = torch.empty(C, dtype=torch.float64, device="cuda")
    g = torch.empty((C, T + 2), dtype=torch.float32, device="cuda")
    _lib.check(_lib.load().nax_potential_grad(eng.h, _lib.ptr(z), T + 2, C, _lib.ptr(pe),
                                              _lib.ptr(g)))
    f = om.sv_nc_potential if noncentred else om.sv_potential
    r64 = r.astype(np.float64)
    for c in range(C):
        zc = Z[c].astype(np.float64)
        u = f(zc, r64, positive="softplus")
        gr = om.central_fd(lambda v: f(v, r64, positive="softplus"), zc, eps=1e-6)
        np.testing.assert_allclose(pe[c].item(), u, rtol=2e-6, atol=1e-3)
        np.testing.assert_allclose(g[c].cpu().numpy(), gr, rtol=2e-4,
                                   atol=2e-5 * (np.abs(gr).max() + 1))


def test_sv_graph_replay_matches_direct_launches_at_fixed_waves_per_chain():
    # The persistent chunks replay as one captured hipGraph; with the waves
    # per chain fixed, direct launches give every draw bit for bit.  The
    # tuned shape is reported.
    from numpyro_amd import random as nrandom
    from numpyro_amd.infer.hmc import Engine
    from numpyro_amd.models import StochasticVolatility
    T, C = 500, 6
    r = _returns(T, 41)
    keys = nrandom.chain_init_keys(nrandom.PRNGKey(2), C)
    z0 = torch.from_numpy(_points(T, C, True, 9))
    eng = Engine(StochasticVolatility(r, noncentred=True), C, "NUTS", 40)
    eng.init(z0, keys)
    eng.run(45)
    assert "waves per chain" in eng.tuning(), eng.tuning()
    for w in (4, 1):
        outs = []
        for graphs in (True, False):
            e = Engine(StochasticVolatility(r, noncentred=True), C, "NUTS", 40)
            e.set_graphs(graphs)
            e.set_launch_shape(w)
            e.init(z0, keys)
            out = e.run(120, collect_fields=("z", "num_steps", "diverging"), size=120)
            outs.append({k: v.cpu().numpy() for k, v in out.items()})
        for k in outs[0]:
            assert np.array_equal(outs[0][k], outs[1][k]), (w, k)
