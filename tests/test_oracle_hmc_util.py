"""Pins the oracle's NUTS/HMC building blocks to the reference's own
known-answer and closed-form tests (test/infer/test_hmc_util.py), re-expressed
against the numpy restatement."""
import math

import numpy as np
from numpy.testing import assert_allclose
import pytest

from oracle import hmc_util as hu
from oracle import jax_random as jr


def test_dual_averaging():
    # test/infer/test_hmc_util.py:36-52
    da_init, da_update = hu.dual_averaging(gamma=0.5)
    st = da_init()
    for _ in range(10):
        x = st[0]
        g = 2 * (x + 1)
        st = da_update(g, st)
    assert_allclose(st[1], -1.0, atol=1e-3)


@pytest.mark.parametrize("diagonal", [True, False])
@pytest.mark.parametrize("regularize", [True, False])
def test_welford_covariance(diagonal, regularize):
    # test/infer/test_hmc_util.py:55-89 plus the exact numpy statistic
    rng = np.random.RandomState(0)
    loc = rng.randn(3)
    a = rng.randn(3, 3)
    target_cov = a @ a.T
    x = rng.multivariate_normal(loc, target_cov, size=(2000,))
    init, update, final = hu.welford_covariance(diagonal=diagonal)
    st = init(3)
    for i in range(2000):
        st = update(x[i], st)
    cov, cov_inv_sqrt, _ = final(st, regularize=regularize)
    n = 2000
    exact = np.var(x, axis=0, ddof=1) if diagonal else np.cov(x.T)
    if regularize:
        exact = (n / (n + 5)) * exact + 1e-3 * 5 / (n + 5) * (1 if diagonal else np.eye(3))
    assert_allclose(cov, exact, rtol=1e-10)
    if diagonal:
        assert_allclose(cov, np.diag(target_cov), rtol=0.06)
        assert_allclose(cov_inv_sqrt, np.sqrt(1 / np.diag(target_cov)), rtol=0.06)
    else:
        assert_allclose(cov, target_cov, rtol=0.06)
        # cov_inv_sqrt is the lower Cholesky factor of the precision
        assert_allclose(cov_inv_sqrt @ cov_inv_sqrt.T, np.linalg.inv(cov), rtol=1e-8)


# ---- velocity Verlet closed forms: test_hmc_util.py:121-229 ------------------

def _sho(q):
    return 0.5 * q[0] ** 2, np.array([q[0]])


def _orbit(q):
    rr = (q[0] ** 2 + q[1] ** 2) ** 0.5
    return -1.0 / rr, np.array([q[0], q[1]]) / rr ** 3


def _quartic(q):
    return 0.25 * q[0] ** 4, np.array([q[0] ** 3])


CASES = [
    (_sho, 0.01, 100, [0.0], [1.0], [np.sin(1.0)], [np.cos(1.0)], 1e-4),
    (_orbit, 0.01, 628, [1.0, 0.0], [0.0, 1.0], [1.0, 0.0], [0.0, 1.0], 5e-3),
    (_quartic, 0.1, 1810, [0.02], [0.0], [-0.02], [0.0], 1e-4),
]


@pytest.mark.parametrize("case", CASES, ids=["sho", "orbit", "quartic"])
def test_velocity_verlet(case):
    f, eps, L, qi, pi, qf, pf, prec = case
    m_inv = np.ones(len(qi))
    init, update = hu.velocity_verlet(f)

    def run(q, p):
        st = init(np.array(q, float), np.array(p, float))
        for _ in range(L):
            st = update(eps, m_inv, st)
        return st

    st = run(qi, pi)
    assert_allclose(st.z, qf, atol=prec)
    assert_allclose(st.r, pf, atol=prec)
    e0 = f(np.array(qi, float))[0] + 0.5 * np.sum(np.square(pi))
    e1 = st.potential_energy + 0.5 * np.sum(st.r ** 2)
    assert_allclose(e0, e1, atol=1e-5)
    back = run(st.z, -st.r)
    assert_allclose(back.z, qi, atol=1e-4)


@pytest.mark.parametrize("init_step_size", [0.1, 10.0])
def test_find_reasonable_step_size(init_step_size):
    # test_hmc_util.py:232-275: threshold (-8 log 0.8)^(1/4)
    gen = lambda proto, m_inv, key: np.array([1.0])  # noqa: E731
    step = hu.find_reasonable_step_size(lambda q: (0.5 * q[0] ** 2, q.copy()), gen,
                                        init_step_size, np.array([1.0]),
                                        (np.array([0.0]), None, None, None), jr.PRNGKey(0))
    threshold = (-math.log(0.8) * 8) ** 0.25
    if init_step_size < threshold:
        assert step / 2 < threshold < step
    else:
        assert step * 2 > threshold > step


@pytest.mark.parametrize("num_steps, expected", [
    (18, [(0, 17)]),
    (50, [(0, 6), (7, 44), (45, 49)]),
    (100, [(0, 14), (15, 89), (90, 99)]),
    (150, [(0, 74), (75, 99), (100, 149)]),
    (200, [(0, 74), (75, 99), (100, 149), (150, 199)]),
    (280, [(0, 74), (75, 99), (100, 229), (230, 279)]),
    # the workloads' schedules
    (500, [(0, 74), (75, 99), (100, 149), (150, 249), (250, 449), (450, 499)]),
    (1000, [(0, 74), (75, 99), (100, 149), (150, 249), (250, 449), (450, 949), (950, 999)]),
])
def test_build_adaptation_schedule(num_steps, expected):
    assert [tuple(w) for w in hu.build_adaptation_schedule(num_steps)] == expected


def test_warmup_adapter():
    # test_hmc_util.py:295-378
    def find_ss(step_size, m_inv, z, key):
        return step_size * 4 if step_size < 1 else step_size / 4

    num_steps = 150
    sched = hu.build_adaptation_schedule(num_steps)
    wa_init, wa_update = hu.warmup_adapter(num_steps, find_ss)
    z = np.ones(3)
    st = wa_init((z, None, None, None), jr.PRNGKey(0), 1.0, mass_matrix_size=3)
    assert st.step_size == 0.25
    assert_allclose(st.inverse_mass_matrix, np.ones(3))
    w = sched[0]
    for t in range(w.start, w.end + 1):
        st = wa_update(t, 0.7 + 0.1 * t / (w.end - w.start), (z,), st)
    last = 0.25
    assert st.window_idx == 1 and st.step_size < last
    assert_allclose(st.inverse_mass_matrix, np.ones(3))
    w = sched[1]
    wl = w.end - w.start
    for t in range(w.start, w.end + 1):
        st = wa_update(t, 0.8 + 0.1 * (t - w.start) / wl, (2 * z,), st)
    assert st.window_idx == 2
    reg = 1e-3 * (5 / (w.end + 1 - w.start + 5))
    assert_allclose(st.inverse_mass_matrix, np.full(3, reg), atol=1e-7)
    imm = st.inverse_mass_matrix
    last = st.step_size
    w = sched[2]
    for t in range(w.start, w.end + 1):
        st = wa_update(t, 0.8, (t * z,), st)
    assert st.window_idx == 3
    assert_allclose(st.step_size, last * 10, atol=1e-6)
    assert_allclose(st.inverse_mass_matrix, imm)


@pytest.mark.parametrize("leaf_idx, ckpt_idxs",
                         [(0, (1, 0)), (6, (3, 2)), (7, (0, 2)), (13, (2, 2)), (15, (0, 3))])
def test_leaf_idx_to_ckpt_idx(leaf_idx, ckpt_idxs):
    assert hu.leaf_idx_to_ckpt_idxs(leaf_idx) == ckpt_idxs


@pytest.mark.parametrize("ckpt_idxs, expected",
                         [((3, 2), False), ((3, 3), True), ((0, 0), False), ((0, 1), True),
                          ((1, 3), True)])
def test_is_iterative_turning(ckpt_idxs, expected):
    imm = np.ones(1)
    r_ckpts = np.array([[1.0], [2.0], [3.0], [-2.0]])
    r_sum_ckpts = np.array([[2.0], [4.0], [4.0], [-1.0]])
    got = hu.is_iterative_turning(imm, np.array([1.0]), np.array([3.0]), r_ckpts,
                                  r_sum_ckpts, *ckpt_idxs)
    assert got == expected


@pytest.mark.parametrize("step_size", [0.01, 1.0, 100.0])
def test_build_tree(step_size):
    # test_hmc_util.py:406-442
    init, update = hu.velocity_verlet(lambda q: (0.5 * q[0] ** 2, q.copy()))
    st = init(np.array([0.0]), np.array([1.0]))
    tree = hu.build_tree(update, st, np.array([1.0]), step_size, jr.PRNGKey(0))
    assert tree.num_proposals >= 2 ** (tree.depth - 1)
    assert tree.sum_accept_probs <= tree.num_proposals
    if tree.depth < 10:
        assert tree.turning or tree.diverging
    if step_size > 10:
        assert tree.diverging and tree.num_proposals == 1
    if step_size < 0.1:
        assert tree.num_proposals > 10
