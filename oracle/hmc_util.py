"""NUTS/HMC building blocks restated in float64 numpy (oracle).

Restates ``numpyro/infer/hmc_util.py`` of numpyro 0.21.0 for a flat latent
vector and a diagonal or dense inverse mass matrix.  Random keys are uint32
pairs from ``oracle.jax_random``; every key split happens in the same place
and order as the reference so that a chain with the same keys draws the same
bits.
"""
from collections import namedtuple
import math

import numpy as np
from scipy.special import expit

from . import jax_random as jr

AdaptWindow = namedtuple("AdaptWindow", ["start", "end"])
HMCAdaptState = namedtuple(
    "HMCAdaptState",
    ["step_size", "inverse_mass_matrix", "mass_matrix_sqrt", "mass_matrix_sqrt_inv",
     "ss_state", "mm_state", "window_idx", "rng_key"])
IntegratorState = namedtuple("IntegratorState", ["z", "r", "potential_energy", "z_grad"])
TreeInfo = namedtuple(
    "TreeInfo",
    ["z_left", "r_left", "z_left_grad", "z_right", "r_right", "z_right_grad",
     "z_proposal", "z_proposal_pe", "z_proposal_grad", "z_proposal_energy",
     "depth", "weight", "r_sum", "turning", "diverging", "sum_accept_probs",
     "num_proposals"])


# ---- dual averaging: hmc_util.py:60-130 -------------------------------------

def dual_averaging(t0=10, kappa=0.75, gamma=0.05):
    def init_fn(prox_center=0.0):
        return (0.0, 0.0, 0.0, 0, prox_center)

    def update_fn(g, state):
        x_t, x_avg, g_avg, t, prox_center = state
        t = t + 1
        g_avg = (1 - 1 / (t + t0)) * g_avg + g / (t + t0)
        x_t = prox_center - (t ** 0.5) / gamma * g_avg
        weight_t = t ** (-kappa)
        x_avg = (1 - weight_t) * x_avg + weight_t * x_t
        return (x_t, x_avg, g_avg, t, prox_center)

    return init_fn, update_fn


# ---- Welford (co)variance: hmc_util.py:133-239 ------------------------------

def welford_covariance(diagonal=True):
    def init_fn(size):
        shape = (size,) if diagonal else (size, size)
        return (np.zeros(size), np.zeros(shape), 0)

    def update_fn(sample, state):
        mean, m2, n = state
        n = n + 1
        delta_pre = sample - mean
        mean = mean + delta_pre / n
        delta_post = sample - mean
        if m2.ndim == 1:
            m2 = m2 + delta_pre * delta_post
        else:
            m2 = m2 + np.outer(delta_post, delta_pre)
        return (mean, m2, n)

    def final_fn(state, regularize=False):
        mean, m2, n = state
        cov = m2 / (n - 1)
        if regularize:
            scaled_cov = (n / (n + 5)) * cov
            shrinkage = 1e-3 * (5 / (n + 5))
            if scaled_cov.ndim == 1:
                cov = scaled_cov + shrinkage
            else:
                cov = scaled_cov + shrinkage * np.identity(mean.shape[0])
        if cov.ndim == 2:
            tril_inv = np.swapaxes(np.linalg.cholesky(cov[::-1, ::-1])[::-1, ::-1], -2, -1)
            cov_inv_sqrt = np.linalg.solve(tril_inv, np.identity(cov.shape[-1]))
        else:
            tril_inv = np.sqrt(cov)
            cov_inv_sqrt = np.reciprocal(tril_inv)
        return cov, cov_inv_sqrt, tril_inv

    return init_fn, update_fn, final_fn


# ---- kinetic energy: hmc_util.py:1183-1223 ----------------------------------

def euclidean_kinetic_energy(inverse_mass_matrix, r):
    v = inverse_mass_matrix @ r if inverse_mass_matrix.ndim == 2 else inverse_mass_matrix * r
    return 0.5 * np.dot(v, r)


def kinetic_grad(inverse_mass_matrix, r):
    return inverse_mass_matrix @ r if inverse_mass_matrix.ndim == 2 else inverse_mass_matrix * r


# ---- velocity Verlet: hmc_util.py:262-311 -----------------------------------

def velocity_verlet(potential_and_grad):
    """``potential_and_grad(z) -> (U, dU/dz)`` replaces jax.value_and_grad."""
    def init_fn(z, r, potential_energy=None, z_grad=None):
        if potential_energy is None or z_grad is None:
            potential_energy, z_grad = potential_and_grad(z)
        return IntegratorState(z, r, potential_energy, z_grad)

    def update_fn(step_size, inverse_mass_matrix, state):
        z, r, _, z_grad = state
        r = r - 0.5 * step_size * z_grad
        r_grad = kinetic_grad(inverse_mass_matrix, r)
        z = z + step_size * r_grad
        potential_energy, z_grad = potential_and_grad(z)
        r = r - 0.5 * step_size * z_grad
        return IntegratorState(z, r, potential_energy, z_grad)

    return init_fn, update_fn


# ---- step-size heuristic: hmc_util.py:314-384 -------------------------------

def find_reasonable_step_size(potential_and_grad, momentum_generator, init_step_size,
                              inverse_mass_matrix, z_info, rng_key):
    target_accept_prob = math.log(0.8)
    _, vv_update = velocity_verlet(potential_and_grad)
    z, _, potential_energy, z_grad = z_info
    if potential_energy is None or z_grad is None:
        potential_energy, z_grad = potential_and_grad(z)
    tiny = np.finfo(np.float64).tiny
    fmax = np.finfo(np.float64).max
    step_size, last_direction, direction = init_step_size, 0, 0
    while True:
        small_ok = (step_size > tiny) or (direction >= 0)
        large_ok = (step_size < fmax) or (direction <= 0)
        if not (small_ok and large_ok and (last_direction == 0 or direction == last_direction)):
            break
        rng_key, rng_key_momentum = jr.split(rng_key)
        step_size = (2.0 ** direction) * step_size
        r = momentum_generator(z, inverse_mass_matrix, rng_key_momentum)
        _, r_new, pe_new, _ = vv_update(step_size, inverse_mass_matrix,
                                        (z, r, potential_energy, z_grad))
        e_cur = euclidean_kinetic_energy(inverse_mass_matrix, r) + potential_energy
        e_new = euclidean_kinetic_energy(inverse_mass_matrix, r_new) + pe_new
        delta = e_new - e_cur
        direction_new = 1 if target_accept_prob < -delta else -1
        last_direction, direction = direction, direction_new
    return step_size


# ---- window schedule: hmc_util.py:387-436 -----------------------------------

def build_adaptation_schedule(num_steps):
    adaptation_schedule = []
    if num_steps < 20:
        adaptation_schedule.append(AdaptWindow(0, num_steps - 1))
        return adaptation_schedule
    start_buffer_size, end_buffer_size, init_window_size = 75, 50, 25
    if (start_buffer_size + end_buffer_size + init_window_size) > num_steps:
        start_buffer_size = int(0.15 * num_steps)
        end_buffer_size = int(0.1 * num_steps)
        init_window_size = num_steps - start_buffer_size - end_buffer_size
    adaptation_schedule.append(AdaptWindow(0, start_buffer_size - 1))
    end_window_start = num_steps - end_buffer_size
    next_window_size = init_window_size
    next_window_start = start_buffer_size
    while next_window_start < end_window_start:
        cur_window_start, cur_window_size = next_window_start, next_window_size
        if 3 * cur_window_size <= end_window_start - cur_window_start:
            next_window_size = 2 * cur_window_size
        else:
            cur_window_size = end_window_start - cur_window_start
        next_window_start = cur_window_start + cur_window_size
        adaptation_schedule.append(AdaptWindow(cur_window_start, next_window_start - 1))
    adaptation_schedule.append(AdaptWindow(end_window_start, num_steps - 1))
    return adaptation_schedule


# ---- mass matrix init: hmc_util.py:439-515 (flat latent) --------------------

def initialize_mass_matrix(size, inverse_mass_matrix, dense_mass):
    if inverse_mass_matrix is None:
        imm = np.identity(size) if dense_mass else np.ones(size)
        return imm, imm.copy(), imm.copy()
    imm = np.asarray(inverse_mass_matrix, dtype=np.float64)
    if dense_mass:
        if imm.ndim == 1:
            imm = np.diag(imm)
        sqrt_inv = np.swapaxes(np.linalg.cholesky(imm[::-1, ::-1])[::-1, ::-1], -2, -1)
        sqrt = np.linalg.solve(sqrt_inv, np.identity(imm.shape[-1]))
        return imm, sqrt, sqrt_inv
    if imm.ndim == 2:
        imm = np.diag(imm)
    sqrt_inv = np.sqrt(imm)
    return imm, np.reciprocal(sqrt_inv), sqrt_inv


# ---- warmup adapter: hmc_util.py:518-707 ------------------------------------

def warmup_adapter(num_adapt_steps, find_reasonable_step_size=None, adapt_step_size=True,
                   adapt_mass_matrix=True, dense_mass=False, target_accept_prob=0.8,
                   regularize_mass_matrix=True):
    if find_reasonable_step_size is None:
        find_reasonable_step_size = lambda step_size, *args: step_size  # noqa: E731
    ss_init, ss_update = dual_averaging()
    mm_init, mm_update, mm_final = welford_covariance(diagonal=not dense_mass)
    adaptation_schedule = build_adaptation_schedule(num_adapt_steps)
    num_windows = len(adaptation_schedule)

    def init_fn(z_info, rng_key, step_size=1.0, inverse_mass_matrix=None,
                mass_matrix_size=None):
        rng_key, rng_key_ss = jr.split(rng_key)
        size = np.size(z_info[0]) if mass_matrix_size is None else mass_matrix_size
        imm, mm_sqrt, mm_sqrt_inv = initialize_mass_matrix(size, inverse_mass_matrix,
                                                           dense_mass)
        if adapt_step_size:
            step_size = find_reasonable_step_size(step_size, imm, z_info, rng_key_ss)
        ss_state = ss_init(math.log(10 * step_size))
        mm_state = mm_init(imm.shape[-1])
        return HMCAdaptState(step_size, imm, mm_sqrt, mm_sqrt_inv, ss_state, mm_state,
                             0, rng_key)

    def _update_at_window_end(z_info, rng_key_ss, state):
        step_size, imm, mm_sqrt, mm_sqrt_inv, ss_state, mm_state, window_idx, rng_key = state
        if adapt_mass_matrix:
            imm, mm_sqrt, mm_sqrt_inv = mm_final(mm_state, regularize=regularize_mass_matrix)
            mm_state = mm_init(imm.shape[-1])
        if adapt_step_size:
            step_size = find_reasonable_step_size(step_size, imm, z_info, rng_key_ss)
            ss_state = ss_init(math.log(10) + math.log(step_size))
        return HMCAdaptState(step_size, imm, mm_sqrt, mm_sqrt_inv, ss_state, mm_state,
                             window_idx, rng_key)

    def update_fn(t, accept_prob, z_info, state):
        step_size, imm, mm_sqrt, mm_sqrt_inv, ss_state, mm_state, window_idx, rng_key = state
        rng_key, rng_key_ss = jr.split(rng_key)
        if adapt_step_size:
            ss_state = ss_update(target_accept_prob - accept_prob, ss_state)
            log_step_size, log_step_size_avg = ss_state[0], ss_state[1]
            step_size = math.exp(log_step_size_avg) if t == num_adapt_steps - 1 \
                else math.exp(log_step_size)
            fi = np.finfo(np.float64)
            step_size = float(np.clip(step_size, fi.tiny, fi.max))
        is_middle_window = (0 < window_idx) and (window_idx < num_windows - 1)
        if adapt_mass_matrix and is_middle_window:
            mm_state = mm_update(z_info[0], mm_state)
        t_at_window_end = t == adaptation_schedule[window_idx][1]
        if t_at_window_end:
            window_idx = window_idx + 1
        state = HMCAdaptState(step_size, imm, mm_sqrt, mm_sqrt_inv, ss_state, mm_state,
                              window_idx, rng_key)
        if t_at_window_end and is_middle_window:
            state = _update_at_window_end(z_info, rng_key_ss, state)
        return state

    return init_fn, update_fn


# ---- NUTS tree: hmc_util.py:710-1180 ----------------------------------------

def _momentum_angle(imm, r_left, r_right, r_sum):
    v_left = kinetic_grad(imm, r_left)
    v_right = kinetic_grad(imm, r_right)
    r_sum = r_sum - (r_left + r_right) / 2
    return np.dot(v_left, r_sum), np.dot(v_right, r_sum)


def is_turning(imm, r_left, r_right, r_sum):
    la, ra = _momentum_angle(imm, r_left, r_right, r_sum)
    return bool((la <= 0) or (ra <= 0))


def _uniform_transition_kernel(current_tree, new_tree):
    return expit(new_tree.weight - current_tree.weight)


def _biased_transition_kernel(current_tree, new_tree):
    p = math.exp(min(new_tree.weight - current_tree.weight, 700.0))
    if new_tree.turning or new_tree.diverging:
        return 0.0
    return min(p, 1.0)


def _combine_tree(current_tree, new_tree, imm, going_right, rng_key, biased_transition):
    if going_right:
        zl, rl, gl = current_tree.z_left, current_tree.r_left, current_tree.z_left_grad
        zr, rr, gr = new_tree.z_right, new_tree.r_right, new_tree.z_right_grad
    else:
        zl, rl, gl = new_tree.z_left, new_tree.r_left, new_tree.z_left_grad
        zr, rr, gr = current_tree.z_right, current_tree.r_right, current_tree.z_right_grad
    r_sum = current_tree.r_sum + new_tree.r_sum
    if biased_transition:
        p = _biased_transition_kernel(current_tree, new_tree)
        turning = new_tree.turning or is_turning(imm, rl, rr, r_sum)
    else:
        p = _uniform_transition_kernel(current_tree, new_tree)
        turning = current_tree.turning
    transition = jr.bernoulli(rng_key, p)
    src = new_tree if transition else current_tree
    return TreeInfo(zl, rl, gl, zr, rr, gr, src.z_proposal, src.z_proposal_pe,
                    src.z_proposal_grad, src.z_proposal_energy, current_tree.depth + 1,
                    np.logaddexp(current_tree.weight, new_tree.weight), r_sum, turning,
                    new_tree.diverging,
                    current_tree.sum_accept_probs + new_tree.sum_accept_probs,
                    current_tree.num_proposals + new_tree.num_proposals)


def _build_basetree(vv_update, z, r, z_grad, imm, step_size, going_right, energy_current,
                    max_delta_energy):
    step_size = step_size if going_right else -step_size
    z_new, r_new, pe_new, g_new = vv_update(step_size, imm, (z, r, energy_current, z_grad))
    energy_new = pe_new + euclidean_kinetic_energy(imm, r_new)
    delta_energy = energy_new - energy_current
    if math.isnan(delta_energy):
        delta_energy = math.inf
    diverging = delta_energy > max_delta_energy
    accept_prob = min(math.exp(-delta_energy) if delta_energy > -700 else math.inf, 1.0)
    return TreeInfo(z_new, r_new, g_new, z_new, r_new, g_new, z_new, pe_new, g_new,
                    energy_new, 0, -delta_energy, r_new, False, diverging, accept_prob, 1)


def leaf_idx_to_ckpt_idxs(n):
    """``hmc_util.py:941-958``."""
    idx_max = bin(n >> 1).count("1")
    num_subtrees = bin((~n & (n + 1)) - 1).count("1")
    idx_min = idx_max - num_subtrees + 1
    return idx_min, idx_max


def is_iterative_turning(imm, r, r_sum, r_ckpts, r_sum_ckpts, idx_min, idx_max):
    """``hmc_util.py:961-981``."""
    i, turning = idx_max, False
    while i >= idx_min and not turning:
        subtree_r_sum = r_sum - r_sum_ckpts[i] + r_ckpts[i]
        turning = is_turning(imm, r_ckpts[i], r, subtree_r_sum)
        i -= 1
    return turning


def _iterative_build_subtree(prototype_tree, vv_update, imm, step_size, going_right,
                             rng_key, energy_current, max_delta_energy, r_ckpts,
                             r_sum_ckpts):
    max_num_proposals = 2 ** prototype_tree.depth
    tree = prototype_tree._replace(num_proposals=0)
    turning = False
    while tree.num_proposals < max_num_proposals and not turning and not tree.diverging:
        rng_key, transition_rng_key = jr.split(rng_key)
        if going_right:
            z, r, g = tree.z_right, tree.r_right, tree.z_right_grad
        else:
            z, r, g = tree.z_left, tree.r_left, tree.z_left_grad
        new_leaf = _build_basetree(vv_update, z, r, g, imm, step_size, going_right,
                                   energy_current, max_delta_energy)
        if tree.num_proposals == 0:
            new_tree = new_leaf
        else:
            new_tree = _combine_tree(tree, new_leaf, imm, going_right, transition_rng_key,
                                     False)
        leaf_idx = tree.num_proposals
        idx_min, idx_max = leaf_idx_to_ckpt_idxs(leaf_idx)
        if leaf_idx % 2 == 0:
            r_ckpts[idx_max] = new_leaf.r_right
            r_sum_ckpts[idx_max] = new_tree.r_sum
        turning = is_iterative_turning(imm, new_leaf.r_right, new_tree.r_sum, r_ckpts,
                                       r_sum_ckpts, idx_min, idx_max)
        tree = new_tree
    return tree._replace(depth=prototype_tree.depth, turning=turning)


def _double_tree(current_tree, vv_update, imm, step_size, going_right, rng_key,
                 energy_current, max_delta_energy, r_ckpts, r_sum_ckpts):
    key, transition_key = jr.split(rng_key)
    new_tree = _iterative_build_subtree(current_tree, vv_update, imm, step_size,
                                        going_right, key, energy_current, max_delta_energy,
                                        r_ckpts, r_sum_ckpts)
    return _combine_tree(current_tree, new_tree, imm, going_right, transition_key, True)


def build_tree(vv_update, verlet_state, imm, step_size, rng_key, max_delta_energy=1000.0,
               max_tree_depth=10):
    """``hmc_util.py:1088-1180``."""
    if isinstance(max_tree_depth, tuple):
        max_tree_depth_current, max_tree_depth = max_tree_depth
    else:
        max_tree_depth_current = max_tree_depth
    z, r, pe, z_grad = verlet_state
    energy_current = pe + euclidean_kinetic_energy(imm, r)
    r_ckpts = np.zeros((max_tree_depth, np.size(r)))
    r_sum_ckpts = np.zeros((max_tree_depth, np.size(r)))
    tree = TreeInfo(z, r, z_grad, z, r, z_grad, z, pe, z_grad, energy_current, 0, 0.0, r,
                    False, False, 0.0, 0)
    key = rng_key
    while tree.depth < max_tree_depth_current and not tree.turning and not tree.diverging:
        keys = jr.split(key, 3)
        key, direction_key, doubling_key = keys[0], keys[1], keys[2]
        going_right = jr.bernoulli(direction_key)
        tree = _double_tree(tree, vv_update, imm, step_size, going_right, doubling_key,
                            energy_current, max_delta_energy, r_ckpts, r_sum_ckpts)
    return tree
