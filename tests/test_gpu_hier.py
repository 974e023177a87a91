"""Ragged hierarchical logistic regressions on the GPU: potential + adjoint
against the float64 oracle, NUTS tree sizes against the oracle, and one
group's posterior against the same model as a torch potential_fn."""
import numpy as np
import pytest
import torch

from oracle import hmc as ohmc
from oracle import jax_random as ojr
from oracle import models as om

pytestmark = pytest.mark.gpu


def _groups(shapes, seed):
    rng = np.random.default_rng(seed)
    Xs, ys = [], []
    for n, d in shapes:
        X = rng.normal(size=(n, d)).astype(np.float32)
        b = rng.normal() + 0.7 * rng.normal(size=d)
        ys.append((rng.uniform(size=n) < 1 / (1 + np.exp(-X @ b))).astype(np.float32))
        Xs.append(X)
    return Xs, ys


SHAPES = [(1, 1), (63, 4), (64, 17), (65, 64), (200, 9), (1000, 33), (5000, 64), (129, 63)]


@pytest.mark.parametrize("K, shape", [(1, None), (4, None), (4, 36), (4, 100), (4, 34),
                                      (1, 97)])
def test_hier_potential_grad(K, shape):
    # shape: chains per workgroup + 32 (waves per chain - 1), e.g. 100 = four
    # chains of four waves, each wave on 64 of a 256-row tile
    from numpyro_amd import _lib
    from numpyro_amd.infer.hmc import Engine
    from numpyro_amd.models import HierarchicalLogistic
    Xs, ys = _groups(SHAPES, K)
    m = HierarchicalLogistic(Xs, ys, chains_per_group=K)
    C = len(SHAPES) * K
    eng = Engine(m, C, "NUTS", 0)
    if shape:
        eng.set_launch_shape(shape)
    rng = np.random.default_rng(5)
    Z = np.zeros((C, m.dim), np.float32)
    for c in range(C):
        d = SHAPES[c // K][1]
        Z[c, :d + 2] = 0.5 * rng.normal(size=d + 2)
    # the group kernel reads latents with the engine's padded stride
    zp = torch.zeros((C, eng_stride(eng)), device="cuda")
    zp[:, :m.dim] = torch.from_numpy(Z).cuda()
    pe = torch.empty(C, dtype=torch.float64, device="cuda")
    g = torch.zeros_like(zp)
    _lib.check(_lib.load().nax_potential_grad(eng.h, _lib.ptr(zp), zp.shape[1], C, _lib.ptr(pe),
                                              _lib.ptr(g)))
    for c in range(C):
        i, d = c // K, SHAPES[c // K][1]
        u, gr = om.hier_potential_grad(Z[c, :d + 2].astype(np.float64), Xs[i], ys[i])
        np.testing.assert_allclose(pe[c].item(), u, rtol=1e-5, atol=1e-4)
        np.testing.assert_allclose(g[c, :d + 2].cpu().numpy(), gr, rtol=1e-4,
                                   atol=1e-5 * (np.abs(gr).max() + 1))


def eng_stride(eng):
    return (eng.D + 63) // 64 * 64


def test_hier_nuts_tree_sizes_match_oracle():
    from numpyro_amd.infer import hmc
    from numpyro_amd.infer.util import fori_collect
    from numpyro_amd.models import HierarchicalLogistic
    shapes = [(150, 3), (90, 6)]
    K = 2
    Xs, ys = _groups(shapes, 9)
    m = HierarchicalLogistic(Xs, ys, chains_per_group=K)
    C = len(shapes) * K
    rng = np.random.default_rng(10)
    Z = np.zeros((C, m.dim), np.float32)
    for c in range(C):
        d = shapes[c // K][1]
        Z[c, :d + 2] = rng.uniform(-1, 1, size=d + 2)
    keys = ojr.split(ojr.PRNGKey(4), C)
    kw = dict(step_size=0.1, adapt_step_size=False, adapt_mass_matrix=False)
    ik, sk = hmc(m)
    st = ik(torch.from_numpy(Z), 0, rng_key=keys, **kw)
    gns, gz = fori_collect(0, 20, sk, st, transform=lambda s: (s.num_steps, s.z))
    gns, gz = gns.cpu().numpy(), gz.cpu().numpy()
    for c in range(C):
        i, d = c // K, shapes[c // K][1]
        oik, osk = ohmc.hmc(lambda q: om.hier_potential_grad(q, Xs[i], ys[i]))
        ost = oik(Z[c, :d + 2].astype(np.float64), 0, rng_key=keys[c], **kw)
        ns, zs = [], []
        for _ in range(20):
            ost = osk(ost)
            ns.append(ost.num_steps)
            zs.append(ost.z)
        assert list(gns[c]) == ns, (c, list(gns[c]), ns)
        np.testing.assert_allclose(gz[c, :, :d + 2], np.stack(zs), rtol=1e-3, atol=1e-3)
        assert np.all(gz[c, :, d + 2:] == 0)


def test_hier_group_posterior_vs_torch_potential():
    from numpyro_amd import random
    from numpyro_amd.diagnostics import effective_sample_size
    from numpyro_amd.infer import MCMC, NUTS
    from numpyro_amd.models import HierarchicalLogistic
    shapes = [(300, 5), (2000, 40), (120, 3), (800, 12)]
    Xs, ys = _groups(shapes, 21)
    K = 16
    m = HierarchicalLogistic(Xs, ys, chains_per_group=K)
    mcmc = MCMC(NUTS(m), num_warmup=300, num_samples=300, num_chains=len(shapes) * K)
    mcmc.run(random.PRNGKey(0), init_params=torch.zeros(len(shapes) * K, m.dim))
    post = mcmc.get_samples(group_by_chain=True)
    mu_h = post["mu"][:K].cpu().numpy().astype(np.float64)
    X0 = torch.from_numpy(Xs[0]).double().cuda()
    y0 = torch.from_numpy(ys[0]).double().cuda()
    D = shapes[0][1]

    def pot(z):
        eta, mu, ut = z[:D], z[D], z[D + 1]
        tau = torch.exp(ut)
        l = X0 @ (mu + tau * eta)
        ll = torch.sum(y0 * l - torch.nn.functional.softplus(l))
        lp = -0.5 * torch.sum(eta ** 2) - 0.5 * mu ** 2 - torch.log1p(tau ** 2) + ut
        return -(ll + lp)

    m2 = MCMC(NUTS(potential_fn=pot), num_warmup=300, num_samples=300, num_chains=K)
    m2.run(random.PRNGKey(1), init_params=torch.zeros(K, D + 2))
    mu_t = m2.get_samples(group_by_chain=True)[..., D].cpu().numpy().astype(np.float64)
    se = np.sqrt(mu_h.var() / effective_sample_size(mu_h) + mu_t.var() / effective_sample_size(mu_t))
    assert abs(mu_h.mean() - mu_t.mean()) < 4 * se, (mu_h.mean(), mu_t.mean(), se)


def test_hier_softplus_potential_grad():
    # tau unconstrained by SoftplusTransform (transforms.py:1156-1175): U
    # against the float64 oracle, dU against central differences of it
    from numpyro_amd import _lib
    from numpyro_amd.infer.hmc import Engine
    from numpyro_amd.models import HierarchicalLogistic
    shapes = [(63, 4), (200, 9), (1000, 33), (65, 64)]
    K = 2
    Xs, ys = _groups(shapes, 17)
    m = HierarchicalLogistic(Xs, ys, chains_per_group=K, positive_transform="softplus")
    C = len(shapes) * K
    eng = Engine(m, C, "NUTS", 0)
    rng = np.random.default_rng(6)
    Z = np.zeros((C, m.dim), np.float32)
    for c in range(C):
        d = shapes[c // K][1]
        Z[c, :d + 2] = 0.5 * rng.normal(size=d + 2)
        Z[c, d + 1] = rng.uniform(-3, 3)
    zp = torch.zeros((C, eng_stride(eng)), device="cuda")
    zp[:, :m.dim] = torch.from_numpy(Z).cuda()
    pe = torch.empty(C, dtype=torch.float64, device="cuda")
    g = torch.zeros_like(zp)
    _lib.check(_lib.load().nax_potential_grad(eng.h, _lib.ptr(zp), zp.shape[1], C, _lib.ptr(pe),
                                              _lib.ptr(g)))
    for c in range(C):
        i, d = c // K, shapes[c // K][1]
        zc = Z[c, :d + 2].astype(np.float64)
        u = om.hier_potential(zc, Xs[i], ys[i], positive="softplus")
        gr = om.central_fd(lambda v: om.hier_potential(v, Xs[i], ys[i], positive="softplus"),
                           zc, eps=1e-6)
        np.testing.assert_allclose(pe[c].item(), u, rtol=1e-5, atol=1e-4)
        np.testing.assert_allclose(g[c, :d + 2].cpu().numpy(), gr, rtol=1e-4,
                                   atol=1e-5 * (np.abs(gr).max() + 1))


def test_hier_graph_replay_matches_direct_launches_and_reports_tuning():
    # nax_run replays the persistent chunks (target step + launch) as one
    # captured hipGraph; direct launches must give every draw bit for bit.
    # The launch shape is (chains per workgroup, waves per chain), code
    # kb + 32 (waves - 1).  The chains-per-workgroup split does not change
    # any draw; the waves per chain set the order of the sums, so runs
    # compare bit for bit within one waves-per-chain setting, and the
    # autotuned shape is kept for the process, so two tuned engines agree.
    from numpyro_amd import random as nrandom
    from numpyro_amd.infer.hmc import Engine
    from numpyro_amd.models import HierarchicalLogistic
    shapes = [(300, 5), (90, 12), (1200, 3), (64, 7)]
    Xs, ys = _groups(shapes, 23)
    K = 4
    C = len(shapes) * K
    keys = nrandom.chain_init_keys(nrandom.PRNGKey(4), C)
    runs = {}
    for graphs, shape in ((True, None), (False, None), (False, 1), (True, 2), (False, 4),
                          (False, 36), (True, 34), (True, 97), (False, 100)):
        eng = Engine(HierarchicalLogistic(Xs, ys, chains_per_group=K), C, "NUTS", 60)
        eng.set_graphs(graphs)
        if shape:
            eng.set_launch_shape(shape)
        eng.init(torch.zeros((C, eng.D)), keys)
        out = eng.run(160, collect_fields=("z", "num_steps", "potential_energy"), size=160)
        tun = eng.tuning()
        assert tun.startswith("hier: ") and "chains per workgroup" in tun, tun
        print(f"[hier graphs={graphs} shape={shape}] {tun}", flush=True)
        runs[(graphs, shape)] = {k: v.cpu().numpy() for k, v in out.items()}
    groups = [[(True, None), (False, None)], [(False, 1), (True, 2), (False, 4)],
              [(False, 36), (True, 34)], [(True, 97), (False, 100)]]
    for grp in groups:
        for other in grp[1:]:
            for k in runs[grp[0]]:
                assert np.array_equal(runs[grp[0]][k], runs[other][k]), (grp[0], other, k)
    # 1, 2 and 4 waves per chain: the same chains up to float rounding of
    # the sums -- equal tree sizes over the first iterations
    for shape in (36, 100):
        assert np.array_equal(runs[(False, 4)]["num_steps"][:, :5],
                              runs[(False, shape)]["num_steps"][:, :5]), shape
        np.testing.assert_allclose(runs[(False, 4)]["z"][:, :3], runs[(False, shape)]["z"][:, :3],
                                   rtol=1e-3, atol=1e-4)


def test_hier_warmup_iterations_follow_oracle_with_batch_keys():
    """One model of the 1024-model workload (make_hier(1024, 4), model 593,
    N=376, D=33) run alone with the keys its chains get in the batch
    (chain_keys(PRNGKey(0), 4096)[4g:4g+4]), from z = 0 with dual averaging
    and diagonal Welford active (500 warmup): the first 8 trees have the
    oracle's sizes and step sizes and the first 6 positions agree within
    1e-3.  Later iterations part ways as fp32 and fp64 round differently
    and an accept decision flips, as any two float implementations do."""
    from numpyro_amd import random as nrandom
    from numpyro_amd.infer.hmc import Engine
    from numpyro_amd.models import HierarchicalLogistic
    from numpyro_amd.workloads import make_hier
    g, K, T = 593, 8, 6
    Xs, ys = make_hier(1024, 4)
    X, y = Xs[g], ys[g]
    D = X.shape[1]
    keys = nrandom.chain_init_keys(nrandom.PRNGKey(0), 4096)[4 * g:4 * g + 4]
    np.testing.assert_array_equal(keys, np.asarray(ohmc.chain_keys(ojr.PRNGKey(0), 4096))[4 * g:4 * g + 4])
    eng = Engine(HierarchicalLogistic([X], [y], chains_per_group=4), 4, "NUTS", 500)
    eng.init(torch.zeros((4, eng.D)), np.asarray(keys, np.uint32))
    out = eng.run(K, collect_fields=("z", "num_steps", "adapt_state.step_size"), start_idx=0,
                  size=K)
    gz = out["z"].cpu().numpy().astype(np.float64)[..., :D + 2]
    gns = out["num_steps"].cpu().numpy()
    gss = out["adapt_state.step_size"].cpu().numpy()
    f = lambda q: om.hier_potential_grad(q, X.astype(np.float64), y.astype(np.float64))
    for c in range(4):
        ik, sk = ohmc.hmc(f)
        st = ik(np.zeros(D + 2), 500, rng_key=keys[c])
        ns, zs, ss = [], [], []
        for _ in range(K):
            st = sk(st)
            ns.append(st.num_steps)
            zs.append(st.z)
            ss.append(st.adapt_state.step_size)
        assert list(gns[c]) == ns, (c, list(gns[c]), ns)
        np.testing.assert_allclose(gss[c], ss, rtol=1e-3)
        np.testing.assert_allclose(gz[c, :T], np.stack(zs)[:T], rtol=1e-3, atol=1e-3)
