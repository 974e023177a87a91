"""Functional HMC/NUTS kernel restated in float64 numpy (oracle).

Restates ``numpyro/infer/hmc.py:85-538`` (``momentum_generator``, ``hmc()``'s
``init_kernel`` / ``sample_kernel`` / ``_nuts_next`` / ``_hmc_next``) for a
flat latent vector, plus the multi-chain key handling of ``MCMC.run``
(``numpyro/infer/mcmc.py:670-671``) and ``HMC.init`` (``hmc.py:744-750``).
"""
from collections import namedtuple
import math

import numpy as np

from . import jax_random as jr
from . import hmc_util as hu

HMCState = namedtuple(
    "HMCState",
    ["i", "z", "z_grad", "potential_energy", "energy", "r", "trajectory_length",
     "num_steps", "accept_prob", "mean_accept_prob", "diverging", "adapt_state",
     "rng_key"])


def momentum_generator(prototype_r, mass_matrix_sqrt, rng_key):
    """``hmc.py:92-110``."""
    eps = jr.normal(rng_key, np.shape(mass_matrix_sqrt)[0])
    if mass_matrix_sqrt.ndim == 1:
        return mass_matrix_sqrt * eps
    return mass_matrix_sqrt @ eps


def hmc(potential_and_grad, algo="NUTS"):
    """``hmc.py:113-538``; ``potential_and_grad(z) -> (U, dU)``."""
    max_delta_energy = 1000.0
    cfg = {}

    def init_kernel(init_params, num_warmup, step_size=1.0, inverse_mass_matrix=None,
                    adapt_step_size=True, adapt_mass_matrix=True, dense_mass=False,
                    target_accept_prob=0.8, *, num_steps=None, trajectory_length=2 * math.pi,
                    max_tree_depth=10, find_heuristic_step_size=False,
                    regularize_mass_matrix=True, rng_key=None):
        rng_key = jr.PRNGKey(0) if rng_key is None else np.asarray(rng_key, np.uint32)
        cfg["wa_steps"] = num_warmup
        cfg["max_treedepth"] = (max_tree_depth if isinstance(max_tree_depth, tuple)
                                else (max_tree_depth, max_tree_depth))
        cfg["fixed_num_steps"] = num_steps
        z = np.asarray(init_params, dtype=np.float64)
        find_ss = None
        if find_heuristic_step_size:
            # hmc.py:319-323 binds momentum_generator, which find_reasonable_step_size
            # calls with the *inverse* mass matrix in the mass_matrix_sqrt slot
            # (hmc_util.py:355); restated as is.
            def find_ss(step_size, imm, z_info, key):
                return hu.find_reasonable_step_size(potential_and_grad, momentum_generator,
                                                    step_size, imm, z_info, key)
        wa_init, wa_update = hu.warmup_adapter(
            num_warmup, find_reasonable_step_size=find_ss, adapt_step_size=adapt_step_size,
            adapt_mass_matrix=adapt_mass_matrix, dense_mass=dense_mass,
            target_accept_prob=target_accept_prob,
            regularize_mass_matrix=regularize_mass_matrix)
        cfg["wa_update"] = wa_update
        keys = jr.split(rng_key, 3)
        rng_key_hmc, rng_key_wa, rng_key_momentum = keys[0], keys[1], keys[2]
        pe, z_grad = potential_and_grad(z)
        z_info = hu.IntegratorState(z, None, pe, z_grad)
        wa_state = wa_init(z_info, rng_key_wa, float(step_size),
                           inverse_mass_matrix=inverse_mass_matrix,
                           mass_matrix_size=np.size(z))
        r = momentum_generator(z, wa_state.mass_matrix_sqrt, rng_key_momentum)
        vv_init, vv_update = hu.velocity_verlet(potential_and_grad)
        cfg["vv_update"] = vv_update
        vv_state = vv_init(z, r, potential_energy=pe, z_grad=z_grad)
        energy = vv_state.potential_energy + hu.euclidean_kinetic_energy(
            wa_state.inverse_mass_matrix, vv_state.r)
        return HMCState(0, vv_state.z, vv_state.z_grad, vv_state.potential_energy, energy,
                        None, trajectory_length, 0, 0.0, 0.0, False, wa_state, rng_key_hmc)

    def _hmc_next(step_size, imm, vv_state, rng_key, trajectory_length):
        vv_update = cfg["vv_update"]
        if cfg["fixed_num_steps"] is not None:
            num_steps = cfg["fixed_num_steps"]
        else:
            num_steps = int(math.ceil(trajectory_length / step_size))
        if trajectory_length is not None:
            step_size = trajectory_length / num_steps
        vv_new = vv_state
        for _ in range(num_steps):
            vv_new = vv_update(step_size, imm, vv_new)
        e_old = vv_state.potential_energy + hu.euclidean_kinetic_energy(imm, vv_state.r)
        e_new = vv_new.potential_energy + hu.euclidean_kinetic_energy(imm, vv_new.r)
        delta = e_new - e_old
        if math.isnan(delta):
            delta = math.inf
        accept_prob = min(math.exp(-delta) if delta > -700 else math.inf, 1.0)
        diverging = delta > max_delta_energy
        if jr.bernoulli(rng_key, accept_prob):
            return vv_new, e_new, num_steps, accept_prob, diverging
        return vv_state, e_old, num_steps, accept_prob, diverging

    def _nuts_next(step_size, imm, vv_state, rng_key, max_treedepth_current):
        tree = hu.build_tree(cfg["vv_update"], vv_state, imm, step_size, rng_key,
                             max_delta_energy=max_delta_energy,
                             max_tree_depth=(max_treedepth_current, max(cfg["max_treedepth"])))
        accept_prob = tree.sum_accept_probs / tree.num_proposals
        vv = hu.IntegratorState(tree.z_proposal, vv_state.r, tree.z_proposal_pe,
                                tree.z_proposal_grad)
        return vv, tree.z_proposal_energy, tree.num_proposals, accept_prob, tree.diverging

    def sample_kernel(hmc_state):
        keys = jr.split(hmc_state.rng_key, 3)
        rng_key, rng_key_momentum, rng_key_transition = keys[0], keys[1], keys[2]
        r = momentum_generator(hmc_state.z, hmc_state.adapt_state.mass_matrix_sqrt,
                               rng_key_momentum)
        vv_state = hu.IntegratorState(hmc_state.z, r, hmc_state.potential_energy,
                                      hmc_state.z_grad)
        wa_steps = cfg["wa_steps"]
        if algo == "HMC":
            length_arg = hmc_state.trajectory_length
            nxt = _hmc_next
        else:
            mt = cfg["max_treedepth"]
            length_arg = mt[0] if hmc_state.i < wa_steps else mt[1]
            nxt = _nuts_next
        vv_state, energy, num_steps, accept_prob, diverging = nxt(
            hmc_state.adapt_state.step_size, hmc_state.adapt_state.inverse_mass_matrix,
            vv_state, rng_key_transition, length_arg)
        adapt_state = hmc_state.adapt_state
        if hmc_state.i < wa_steps:
            adapt_state = cfg["wa_update"](hmc_state.i, accept_prob, vv_state, adapt_state)
        itr = hmc_state.i + 1
        n = itr if hmc_state.i < wa_steps else itr - wa_steps
        mean_accept_prob = hmc_state.mean_accept_prob + \
            (accept_prob - hmc_state.mean_accept_prob) / n
        return HMCState(itr, vv_state.z, vv_state.z_grad, vv_state.potential_energy, energy,
                        None, hmc_state.trajectory_length, num_steps, accept_prob,
                        mean_accept_prob, diverging, adapt_state, rng_key)

    return init_kernel, sample_kernel


def chain_keys(rng_key, num_chains):
    """Per-chain ``init_kernel`` keys as ``MCMC.run`` + ``HMC.init`` derive
    them: split(key, C) (``mcmc.py:670-671``), then split(k_c)[0]
    (``hmc.py:744-750``)."""
    keys = jr.split(rng_key, num_chains) if num_chains > 1 else np.asarray(rng_key)[None]
    return np.stack([jr.split(k)[0] for k in keys])


def run_chains(potential_and_grad, init_params, rng_key, num_warmup, num_samples,
               algo="NUTS", collect=("z", "num_steps", "accept_prob", "diverging",
                                     "potential_energy", "step_size"), **init_kwargs):
    """Runs C independent chains (the ``vectorized`` chain method) and returns
    a dict of [C, num_samples, ...] arrays of the post-warmup states."""
    init_params = np.atleast_2d(np.asarray(init_params, dtype=np.float64))
    C = init_params.shape[0]
    keys = chain_keys(rng_key, C)
    out = {k: [] for k in collect}
    for c in range(C):
        init_kernel, sample_kernel = hmc(potential_and_grad, algo=algo)
        st = init_kernel(init_params[c], num_warmup, rng_key=keys[c], **init_kwargs)
        rows = {k: [] for k in collect}
        for it in range(num_warmup + num_samples):
            st = sample_kernel(st)
            if it >= num_warmup:
                for k in collect:
                    v = st.adapt_state.step_size if k == "step_size" else getattr(st, k)
                    rows[k].append(np.copy(v))
        for k in collect:
            out[k].append(np.stack(rows[k]))
    return {k: np.stack(v) for k, v in out.items()}
