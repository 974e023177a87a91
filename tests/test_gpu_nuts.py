"""Whole-chain parity of the GPU engine against the float64 oracle.

With identical per-chain keys, the engine draws the same threefry bits as
the oracle, so every doubling direction, transition and U-turn decision is
the same unless float32 rounding moves a comparison across its threshold.
Without adaptation every decision is far from its threshold on these
problems and whole chains must match (tree sizes exactly, positions to
float32 accuracy).  With adaptation, the early dual-averaging steps amplify
rounding: the float64 oracle run against itself with its potential
evaluated at float32-rounded positions drifts to ~4e-4 after 40 warmup
iterations, enough to flip a near-tie multinomial transition later on.
There the tree sizes must still match for every iteration, positions must
match over the first 25 iterations."""
import numpy as np
import pytest
import torch

from oracle import hmc as ohmc
from oracle import jax_random as ojr
from oracle import models as om

pytestmark = pytest.mark.gpu


def _problem(N, D, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(N, D)).astype(np.float32)
    beta = rng.normal(size=D)
    y = (rng.uniform(size=N) < 1 / (1 + np.exp(-X @ beta))).astype(np.float32)
    return X, y


def _run_engine(X, y, z0, keys, num_warmup, num_samples, algo="NUTS", model_kw=None, **kw):
    from numpyro_amd.infer import hmc
    from numpyro_amd.infer.util import fori_collect
    from numpyro_amd.models import LogisticRegression
    m = LogisticRegression(X, y, **(model_kw or {"kernel": "tile"}))
    init_kernel, sample_kernel = hmc(m, algo=algo)
    st = init_kernel(torch.from_numpy(z0), num_warmup, rng_key=keys, **kw)
    out = fori_collect(0, num_warmup + num_samples, sample_kernel, st,
                       transform=lambda s: (s.z, s.num_steps, s.potential_energy,
                                            s.accept_prob))
    return [o.cpu().numpy() for o in out]


def _run_oracle(X, y, z0, keys, num_warmup, num_samples, algo="NUTS", **kw):
    f = lambda q: om.logistic_potential_grad(q, X, y, 1.0)  # noqa: E731
    C = z0.shape[0]
    res = {"z": [], "num_steps": [], "potential_energy": [], "accept_prob": []}
    for c in range(C):
        ik, sk = ohmc.hmc(f, algo=algo)
        st = ik(z0[c].astype(np.float64), num_warmup, rng_key=keys[c], **kw)
        rows = {k: [] for k in res}
        for _ in range(num_warmup + num_samples):
            st = sk(st)
            for k in res:
                rows[k].append(np.copy(getattr(st, k)))
        for k in res:
            res[k].append(np.stack(rows[k]))
    return [np.stack(res[k]) for k in ("z", "num_steps", "potential_energy", "accept_prob")]


@pytest.mark.parametrize("kernel", ["tile", "chain"])
def test_deep_trees_match_oracle(kernel):
    # a small fixed step: trees of 32-1023 leaves, so a leaf's U-turn checks
    # run against up to depth checkpoints (three in the leaf's fused pass,
    # the rest one by one); tick kernel ("tile") and persistent chain group
    N, D, C = 200, 5, 3
    X, y = _problem(N, D, 31)
    rng = np.random.default_rng(131)
    z0 = rng.uniform(-1, 1, size=(C, D)).astype(np.float32)
    keys = np.stack([ojr.split(ojr.PRNGKey(31), C)[c] for c in range(C)])
    kw = dict(adapt_step_size=False, adapt_mass_matrix=False, step_size=0.01)
    gz, gns, gpe, _ = _run_engine(X, y, z0, keys, 0, 12, model_kw={"kernel": kernel}, **kw)
    oz, ons, ope, _ = _run_oracle(X, y, z0, keys, 0, 12, **kw)
    assert ons.max() >= 31, ons
    assert np.array_equal(gns, ons), (gns, ons)
    np.testing.assert_allclose(gz, oz, rtol=1e-3, atol=1e-3)
    np.testing.assert_allclose(gpe, ope, rtol=1e-5)


@pytest.mark.parametrize("seed, adapt, regularize", [(11, True, True), (12, False, True),
                                                     (13, True, False)])
def test_nuts_chain_matches_oracle(seed, adapt, regularize):
    N, D, C = 200, 5, 3
    X, y = _problem(N, D, seed)
    rng = np.random.default_rng(seed + 100)
    z0 = rng.uniform(-2, 2, size=(C, D)).astype(np.float32)
    keys = np.stack([ojr.split(ojr.PRNGKey(seed), C)[c] for c in range(C)])
    kw = dict(adapt_step_size=adapt, adapt_mass_matrix=adapt,
              step_size=1.0 if adapt else 0.2, regularize_mass_matrix=regularize)
    num_warmup, num_samples = 40, 20
    gz, gns, gpe, gap = _run_engine(X, y, z0, keys, num_warmup, num_samples, **kw)
    oz, ons, ope, oap = _run_oracle(X, y, z0, keys, num_warmup, num_samples, **kw)
    assert np.array_equal(gns, ons), (gns, ons)
    horizon = 25 if adapt else num_warmup + num_samples
    np.testing.assert_allclose(gz[:, :horizon], oz[:, :horizon], rtol=1e-3, atol=1e-3)
    np.testing.assert_allclose(gpe[:, :horizon], ope[:, :horizon], rtol=1e-4 if adapt else 1e-5)
    np.testing.assert_allclose(gap[:, :horizon], oap[:, :horizon], rtol=1e-3, atol=1e-4)


def test_hmc_chain_matches_oracle():
    N, D, C = 150, 4, 2
    X, y = _problem(N, D, 21)
    z0 = np.random.default_rng(3).uniform(-2, 2, size=(C, D)).astype(np.float32)
    keys = np.stack([ojr.split(ojr.PRNGKey(5), C)[c] for c in range(C)])
    kw = dict(trajectory_length=1.0, step_size=0.1)
    gz, gns, gpe, gap = _run_engine(X, y, z0, keys, 30, 20, algo="HMC", **kw)
    oz, ons, ope, oap = _run_oracle(X, y, z0, keys, 30, 20, algo="HMC", **kw)
    assert np.array_equal(gns, ons)
    np.testing.assert_allclose(gz, oz, rtol=1e-3, atol=1e-3)
    np.testing.assert_allclose(gpe, ope, rtol=1e-5)


def _step_sizes(run, X, y, z0, keys, num_warmup, num_samples, **kw):
    """adapt_state.step_size after init and after every iteration, [C, 1 + S]."""
    if run == "engine":
        from numpyro_amd.infer import hmc
        from numpyro_amd.models import LogisticRegression
        init_kernel, sample_kernel = hmc(LogisticRegression(X, y, kernel="tile"), algo="NUTS")
        st = init_kernel(torch.from_numpy(z0), num_warmup, rng_key=keys, **kw)
        rows = [torch.as_tensor(st.adapt_state.step_size).cpu().numpy().reshape(-1)]
        ns = []
        for _ in range(num_warmup + num_samples):
            st = sample_kernel(st)
            rows.append(torch.as_tensor(st.adapt_state.step_size).cpu().numpy().reshape(-1))
            ns.append(torch.as_tensor(st.num_steps).cpu().numpy().reshape(-1))
        return np.stack(rows, 1).astype(np.float64), np.stack(ns, 1)
    f = lambda q: om.logistic_potential_grad(q, X, y, 1.0)  # noqa: E731
    ss, ns = [], []
    for c in range(z0.shape[0]):
        ik, sk = ohmc.hmc(f, algo="NUTS")
        st = ik(z0[c].astype(np.float64), num_warmup, rng_key=keys[c], **kw)
        r, n = [float(st.adapt_state.step_size)], []
        for _ in range(num_warmup + num_samples):
            st = sk(st)
            r.append(float(st.adapt_state.step_size))
            n.append(int(st.num_steps))
        ss.append(r)
        ns.append(n)
    return np.array(ss), np.array(ns)


@pytest.mark.parametrize("init_step", [1.0, 1e-3])
def test_find_heuristic_step_size_matches_oracle(init_step):
    # find_reasonable_step_size (hmc_util.py:314-384) runs at init and at the
    # end of the one middle window of a 60-step warmup ([0,8], [9,53], [54,59]).
    N, D, C = 200, 5, 3
    X, y = _problem(N, D, 31)
    z0 = np.random.default_rng(7).uniform(-2, 2, size=(C, D)).astype(np.float32)
    keys = np.stack([ojr.split(ojr.PRNGKey(17), C)[c] for c in range(C)])
    kw = dict(step_size=init_step, find_heuristic_step_size=True)
    gss, gns = _step_sizes("engine", X, y, z0, keys, 60, 10, **kw)
    oss, ons = _step_sizes("oracle", X, y, z0, keys, 60, 10, **kw)
    # the init heuristic multiplies init_step by a power of two: exact
    np.testing.assert_array_equal(gss[:, 0], oss[:, 0])
    assert np.all(gss[:, 0] != init_step)
    # Dual averaging amplifies float32 rounding (the oracle against itself
    # drifts ~1e-3 by the end of warmup) until a near-tie transition may
    # flip; up to that point tree sizes are identical and step sizes agree
    # (a wrong doubling or halving at the window end would be a factor of 2).
    same = (gns == ons).all(axis=0)
    agree = int(np.argmin(same)) if not same.all() else gns.shape[1]
    assert agree >= 40, (agree, gns, ons)
    np.testing.assert_allclose(gss[:, :26], oss[:, :26], rtol=1e-3)
    np.testing.assert_allclose(gss[:, :41], oss[:, :41], rtol=2e-2)
    # with init_step 1.0 the chains agree through the window-end heuristic
    # (iteration 53) to the end
    if init_step == 1.0:
        assert agree == gns.shape[1]
        np.testing.assert_allclose(gss, oss, rtol=2e-2)


def _spd(D, seed):
    a = np.random.default_rng(seed).normal(size=(D, D))
    return (a @ a.T / D + 0.5 * np.eye(D)).astype(np.float32)


def _run_pair(X, y, z0, keys, num_warmup, num_samples, **kw):
    """Engine and oracle chains side by side: per-iteration z and num_steps,
    and the final adapt_state (inverse mass matrix, mass_matrix_sqrt)."""
    from numpyro_amd.infer import hmc
    from numpyro_amd.models import LogisticRegression
    init_kernel, sample_kernel = hmc(LogisticRegression(X, y, kernel="tile"), algo="NUTS")
    st = init_kernel(torch.from_numpy(z0), num_warmup, rng_key=keys, **kw)
    gz, gns = [], []
    for _ in range(num_warmup + num_samples):
        st = sample_kernel(st)
        gz.append(torch.as_tensor(st.z).cpu().numpy())
        gns.append(torch.as_tensor(st.num_steps).cpu().numpy().reshape(-1))
    g = dict(z=np.stack(gz, 1), ns=np.stack(gns, 1),
             imm=torch.as_tensor(st.adapt_state.inverse_mass_matrix).cpu().numpy(),
             msq=torch.as_tensor(st.adapt_state.mass_matrix_sqrt).cpu().numpy())
    f = lambda q: om.logistic_potential_grad(q, X, y, 1.0)  # noqa: E731
    o = dict(z=[], ns=[], imm=[], msq=[])
    for c in range(z0.shape[0]):
        ik, sk = ohmc.hmc(f, algo="NUTS")
        s = ik(z0[c].astype(np.float64), num_warmup, rng_key=keys[c], **kw)
        zz, nn = [], []
        for _ in range(num_warmup + num_samples):
            s = sk(s)
            zz.append(np.copy(s.z))
            nn.append(int(s.num_steps))
        o["z"].append(zz)
        o["ns"].append(nn)
        o["imm"].append(s.adapt_state.inverse_mass_matrix)
        o["msq"].append(s.adapt_state.mass_matrix_sqrt)
    return g, {k: np.array(v) for k, v in o.items()}


def test_dense_mass_fixed_matrix_matches_oracle():
    # dense inverse mass matrix given, no adaptation: the initial
    # factorisation (hmc_util.py:490-507) and every dense product must agree
    N, D, C = 200, 4, 2
    X, y = _problem(N, D, 41)
    z0 = np.random.default_rng(9).uniform(-2, 2, size=(C, D)).astype(np.float32)
    keys = np.stack([ojr.split(ojr.PRNGKey(23), C)[c] for c in range(C)])
    imm = _spd(D, 5)
    kw = dict(dense_mass=True, adapt_step_size=False, adapt_mass_matrix=False,
              step_size=0.3, inverse_mass_matrix=imm)
    g, o = _run_pair(X, y, z0, keys, 10, 30, **kw)
    np.testing.assert_allclose(g["msq"], o["msq"], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(g["imm"], o["imm"], rtol=1e-6)
    assert np.array_equal(g["ns"], o["ns"]), (g["ns"], o["ns"])
    np.testing.assert_allclose(g["z"], o["z"], rtol=1e-3, atol=1e-3)


def test_dense_mass_adaptation_matches_oracle():
    # dense Welford over the middle window [9, 53] of a 60-step warmup, the
    # regularised covariance and its factors (hmc_util.py:133-239)
    N, D, C = 200, 4, 2
    X, y = _problem(N, D, 43)
    z0 = np.random.default_rng(10).uniform(-2, 2, size=(C, D)).astype(np.float32)
    keys = np.stack([ojr.split(ojr.PRNGKey(29), C)[c] for c in range(C)])
    g, o = _run_pair(X, y, z0, keys, 60, 10, dense_mass=True)
    # adapted chains agree until float32 rounding, amplified by dual
    # averaging, flips a near-tie transition (see the module docstring); a
    # chain that agrees to the end went through the dense window end (53)
    n = g["ns"].shape[1]
    agree = [int(np.argmin(g["ns"][c] == o["ns"][c])) if (g["ns"][c] != o["ns"][c]).any()
             else n for c in range(C)]
    assert min(agree) >= 25, (agree, g["ns"], o["ns"])
    np.testing.assert_allclose(g["z"][:, :15], o["z"][:, :15], rtol=1e-3, atol=1e-3)
    full = [c for c in range(C) if agree[c] == n]
    np.testing.assert_allclose(g["z"][full], o["z"][full], rtol=1e-2, atol=1e-2)
    assert full, agree
    np.testing.assert_allclose(g["imm"][full], o["imm"][full], rtol=2e-2, atol=1e-4)
    np.testing.assert_allclose(g["msq"][full], o["msq"][full], rtol=2e-2, atol=1e-3)
    # S S^T = M = (M^-1)^-1 for the engine's own factors
    for c in range(C):
        m = np.linalg.inv(g["imm"][c].astype(np.float64))
        np.testing.assert_allclose(g["msq"][c] @ g["msq"][c].T, m, rtol=1e-3, atol=1e-3)
        assert np.allclose(np.triu(g["msq"][c], 1), 0)


def test_compacted_mfma_chains_match_oracle():
    # 160 chains on the fp32 MFMA kernel: two chain tiles, and from the first
    # chain to finish an iteration batch on, each tick evaluates only the
    # compacted list of chains that asked for a gradient
    N, D, C = 200, 5, 160
    X, y = _problem(N, D, 51)
    z0 = np.random.default_rng(12).uniform(-2, 2, size=(C, D)).astype(np.float32)
    keys = np.stack([ojr.split(ojr.PRNGKey(31), C)[c] for c in range(C)])
    kw = dict(adapt_step_size=False, adapt_mass_matrix=False, step_size=0.2)
    gz, gns, gpe, _ = _run_engine(X, y, z0, keys, 0, 12, model_kw=dict(kernel="mfma"), **kw)
    oz, ons, ope, _ = _run_oracle(X, y, z0, keys, 0, 12, **kw)
    assert np.array_equal(gns, ons), np.argwhere(gns != ons)[:5]
    np.testing.assert_allclose(gz, oz, rtol=1e-3, atol=1e-3)
    np.testing.assert_allclose(gpe, ope, rtol=1e-5)


def test_chain_kernel_matches_oracle():
    # one wavefront per chain, whole loop in one persistent launch (the
    # "auto" choice for a handful of chains on a small design)
    N, D, C = 300, 6, 4
    X, y = _problem(N, D, 61)
    z0 = np.random.default_rng(14).uniform(-2, 2, size=(C, D)).astype(np.float32)
    keys = np.stack([ojr.split(ojr.PRNGKey(37), C)[c] for c in range(C)])
    kw = dict(adapt_step_size=False, adapt_mass_matrix=False, step_size=0.2)
    gz, gns, gpe, gap = _run_engine(X, y, z0, keys, 0, 40, model_kw=dict(kernel="chain"), **kw)
    oz, ons, ope, oap = _run_oracle(X, y, z0, keys, 0, 40, **kw)
    assert np.array_equal(gns, ons), (gns, ons)
    np.testing.assert_allclose(gz, oz, rtol=1e-3, atol=1e-3)
    np.testing.assert_allclose(gpe, ope, rtol=1e-5)
    np.testing.assert_allclose(gap, oap, rtol=1e-3, atol=1e-4)


def test_chain_kernel_adapted_matches_oracle():
    N, D, C = 300, 6, 4
    X, y = _problem(N, D, 62)
    z0 = np.random.default_rng(15).uniform(-2, 2, size=(C, D)).astype(np.float32)
    keys = np.stack([ojr.split(ojr.PRNGKey(41), C)[c] for c in range(C)])
    gz, gns, gpe, _ = _run_engine(X, y, z0, keys, 40, 20, model_kw=dict(kernel="chain"))
    oz, ons, ope, _ = _run_oracle(X, y, z0, keys, 40, 20)
    # adapted chains agree until amplified float32 rounding flips a near tie
    for c in range(C):
        bad = np.flatnonzero(gns[c] != ons[c])
        assert bad.size == 0 or bad[0] >= 25, (c, gns[c], ons[c])
    np.testing.assert_allclose(gz[:, :15], oz[:, :15], rtol=1e-3, atol=1e-3)


@pytest.mark.parametrize("waves", [1, 2, 8, 16])
def test_chain_kernel_fixed_waves_match_oracle(waves):
    # the per-chain form at a fixed number of waves per chain: one wave, or a
    # group whose first wave runs the state machine alone while the whole
    # group evaluates the potential; graph replay and direct launches give
    # the same draws, and the draws follow the oracle
    from numpyro_amd.infer.hmc import Engine
    from numpyro_amd.models import LogisticRegression
    N, D, C = 300, 6, 4
    X, y = _problem(N, D, 63)
    z0 = np.random.default_rng(16).uniform(-2, 2, size=(C, D)).astype(np.float32)
    keys = np.stack([ojr.split(ojr.PRNGKey(43), C)[c] for c in range(C)])
    kw = dict(adapt_step_size=False, adapt_mass_matrix=False, step_size=0.2)
    outs = []
    for graphs in (True, False):
        e = Engine(LogisticRegression(X, y, kernel="chain"), C, "NUTS", 0, **kw)
        e.set_graphs(graphs)
        e.set_launch_shape(waves)
        e.init(torch.from_numpy(z0), keys)
        out = e.run(40, collect_fields=("z", "num_steps"), size=40)
        outs.append({k: v.cpu().numpy() for k, v in out.items()})
        assert f"{waves} waves per chain" in e.tuning(), e.tuning()
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k
    oz, ons, _, _ = _run_oracle(X, y, z0, keys, 0, 40, **kw)
    assert np.array_equal(outs[0]["num_steps"], ons), (outs[0]["num_steps"], ons)
    np.testing.assert_allclose(outs[0]["z"], oz, rtol=1e-3, atol=1e-3)
