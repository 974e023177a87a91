"""HMC / NUTS kernels with the reference's kernel API, running on the engine.

Mirrors ``numpyro/infer/hmc.py`` (numpyro 0.21.0): ``hmc()`` returns
``(init_kernel, sample_kernel)`` with the same keyword arguments and
defaults (``hmc.py:113-538``); ``HMC`` and ``NUTS`` implement ``MCMCKernel``
(``hmc.py:541-951``).  States are ``HMCState`` namedtuples whose leaves are
torch tensors with a leading chain dimension.  All numerical work (leapfrog,
tree building, adaptation, RNG, potential and gradient) runs in the HIP
library; this module only marshals arguments.
"""
from collections import namedtuple
import ctypes as C
import math
import warnings

import numpy as np
import torch

from .. import _lib
from .. import random as nrandom
from ..compile import compile_model, is_model_function
from ..models import Model, PotentialFn
from .mcmc import MCMCKernel

HMCState = namedtuple(
    "HMCState",
    ["i", "z", "z_grad", "potential_energy", "energy", "r", "trajectory_length",
     "num_steps", "accept_prob", "mean_accept_prob", "diverging", "adapt_state",
     "rng_key"])
HMCAdaptState = namedtuple(
    "HMCAdaptState",
    ["step_size", "inverse_mass_matrix", "mass_matrix_sqrt", "mass_matrix_sqrt_inv",
     "ss_state", "mm_state", "window_idx", "rng_key"])

FIELD_DTYPES = {
    "potential_energy": torch.float64, "energy": torch.float64,
    "accept_prob": torch.float32, "mean_accept_prob": torch.float32,
    "num_steps": torch.int32, "diverging": torch.uint8,
    "adapt_state.step_size": torch.float32,
}
_COLLECT_PTR = {
    "potential_energy": "d_potential_energy", "energy": "d_energy",
    "accept_prob": "d_accept_prob", "mean_accept_prob": "d_mean_accept_prob",
    "num_steps": "d_num_steps", "diverging": "d_diverging",
    "adapt_state.step_size": "d_step_size",
}


class Engine:
    """One sampler handle: C chains of one potential on the current GPU."""

    def __init__(self, model, num_chains, algo, num_warmup, step_size=1.0,
                 adapt_step_size=True, adapt_mass_matrix=True, dense_mass=False,
                 target_accept_prob=0.8, num_steps=None, trajectory_length=2 * math.pi,
                 max_tree_depth=10, find_heuristic_step_size=False,
                 regularize_mass_matrix=True, partitionable=None):
        if algo not in ("HMC", "NUTS"):
            raise ValueError("`algo` must be one of `HMC` or `NUTS`.")
        if isinstance(dense_mass, (list, tuple)):
            if dense_mass:
                raise ValueError("block-structured dense_mass (a list of site groups) is not "
                                 "supported; use dense_mass=True for a full matrix")
            dense_mass = False
        self.dense = bool(dense_mass)
        if not torch.cuda.is_available():
            raise RuntimeError("numpyro_amd needs a HIP GPU (torch.cuda.is_available() is False)")
        self.lib = _lib.load()
        self.model = model
        self.C = int(num_chains)
        self.D = int(model.dim)
        self.algo = algo
        self.num_warmup = int(num_warmup)
        mtd = max_tree_depth if isinstance(max_tree_depth, tuple) else (max_tree_depth,) * 2
        cfg = _lib.nax_config()
        cfg.num_chains = self.C
        cfg.dim = self.D
        cfg.algo = _lib.NAX_ALGO_NUTS if algo == "NUTS" else _lib.NAX_ALGO_HMC
        cfg.num_warmup = self.num_warmup
        cfg.max_tree_depth_warmup = int(mtd[0])
        cfg.max_tree_depth = int(mtd[1])
        cfg.adapt_step_size = int(bool(adapt_step_size))
        cfg.adapt_mass_matrix = int(bool(adapt_mass_matrix))
        cfg.regularize_mass_matrix = int(bool(regularize_mass_matrix))
        part = nrandom.PARTITIONABLE if partitionable is None else partitionable
        self.partitionable = bool(part)
        cfg.partitionable_keys = int(self.partitionable)
        cfg.num_steps = int(num_steps) if num_steps is not None else 0
        cfg.find_heuristic_step_size = int(bool(find_heuristic_step_size))
        cfg.dense_mass = int(self.dense)
        cfg.step_size = float(step_size)
        cfg.target_accept_prob = float(target_accept_prob)
        cfg.trajectory_length = float(trajectory_length) if trajectory_length is not None else 0.0
        cfg.max_delta_energy = 1000.0
        self.cfg = cfg
        h = C.c_void_p()
        _lib.check(self.lib.nax_create(C.byref(cfg), C.byref(h)), "nax_create")
        self.h = h
        model.attach(h)
        self.external = isinstance(model, PotentialFn)
        self.grad_evals = 0
        # HMCECS: a Gibbs update of the subsample before every iteration
        self.gibbs = hasattr(model, "num_blocks")
        self.gibbs_accept = (torch.zeros(self.C, dtype=torch.float32, device="cuda")
                             if self.gibbs else None)

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self.lib.nax_destroy(h)
            self.h = None

    # -- external potential exchange ------------------------------------------
    def _serve_external(self):
        """Evaluates the torch potential at the engine's requested positions
        and hands U, dU back (``nax_exchange``)."""
        zt = torch.empty((self.C, self.D), dtype=torch.float32, device="cuda")
        _lib.check(self.lib.nax_exchange(self.h, _lib.ptr(zt), None, None), "nax_exchange")
        u, g = self.model.value_and_grad(zt)
        u = u.to(torch.float64).contiguous()
        g = g.to(torch.float32).contiguous()
        torch.cuda.synchronize()
        _lib.check(self.lib.nax_exchange(self.h, None, _lib.ptr(u), _lib.ptr(g)), "nax_exchange")

    def _advance_external(self, target, col):
        pending = C.c_int32()
        evals = 0
        while True:
            _lib.check(self.lib.nax_advance(self.h, int(target), col, C.byref(pending)),
                       "nax_advance")
            if pending.value == 0:
                return evals
            evals += pending.value
            self._serve_external()

    # -- init / run -------------------------------------------------------------
    def init(self, z0, keys, inverse_mass_matrix=None):
        z0 = torch.as_tensor(z0, dtype=torch.float32).to("cuda").reshape(self.C, self.D)
        z0 = z0.contiguous()
        keys = np.ascontiguousarray(np.asarray(keys, dtype=np.uint32).reshape(self.C, 2))
        imm = None
        if inverse_mass_matrix is not None:
            imm = torch.as_tensor(inverse_mass_matrix, dtype=torch.float32).to("cuda")
            if self.dense:  # _initialize_mass_matrix: a vector becomes diag(v)
                imm = torch.diag(imm) if imm.ndim == 1 else imm
                imm = imm.reshape(self.D, self.D).contiguous()
            else:
                if imm.ndim == 2:
                    imm = torch.diagonal(imm)
                imm = imm.reshape(self.D).contiguous()
        torch.cuda.synchronize()
        _lib.check(self.lib.nax_init(self.h, _lib.ptr(z0), keys.ctypes.data_as(C.c_void_p),
                                     _lib.ptr(imm)), "nax_init")
        if self.external:
            self._serve_external()
            self._advance_external(0, None)
        self._z0 = z0

    def run(self, target, collect_fields=(), start_idx=0, thinning=1, size=0):
        """Advance every chain to `target` total iterations; returns the
        collected [C, size, ...] tensors."""
        out = {}
        col = _lib.nax_collect()
        col.start_idx = int(start_idx)
        col.thinning = int(thinning)
        col.size = int(size)
        if size > 0:
            for f in collect_fields:
                if f == "_gibbs_accept_prob" and self.gibbs:
                    continue
                if f == "z":
                    out[f] = torch.zeros((self.C, size, self.D), dtype=torch.float32, device="cuda")
                    col.d_z = out[f].data_ptr()
                elif f in _COLLECT_PTR:
                    out[f] = torch.zeros((self.C, size), dtype=FIELD_DTYPES[f], device="cuda")
                    setattr(col, _COLLECT_PTR[f], out[f].data_ptr())
                else:
                    raise ValueError(f"cannot collect field {f!r}")
        gib = gacc = None
        if self.gibbs and size > 0 and "z" in collect_fields:
            # the subsample of each collected iteration (HMCECS's Gibbs site)
            gib = torch.zeros((self.C, size, self.model.m), dtype=torch.int32, device="cuda")
            out["_gibbs"] = gib
        if self.gibbs and size > 0 and "_gibbs_accept_prob" in collect_fields:
            # HMCECSState.accept_prob of each collected iteration
            gacc = torch.zeros((self.C, size), dtype=torch.float32, device="cuda")
            out["_gibbs_accept_prob"] = gacc
        torch.cuda.synchronize()
        if self.external:
            self.grad_evals = self._advance_external(target, C.byref(col))
        elif self.gibbs:
            # HMCECS.sample (numpyro/infer/hmc_gibbs.py:640-681): Gibbs update
            # of the subsample, then one iteration of the inner kernel
            total = 0
            for t in range(int(self.get_state().i.max()), int(target)):
                _lib.check(self.lib.nax_ecs_gibbs(self.h, _lib.ptr(self.gibbs_accept)),
                           "nax_ecs_gibbs")
                ge = C.c_int64()
                _lib.check(self.lib.nax_run(self.h, t + 1, C.byref(col), C.byref(ge)), "nax_run")
                total += ge.value
                rel = t - int(start_idx)
                if rel >= 0 and rel % thinning == thinning - 1 and rel // thinning < size:
                    if gib is not None:
                        gib[:, rel // thinning].copy_(self.model.idx)
                    if gacc is not None:
                        gacc[:, rel // thinning].copy_(self.gibbs_accept)
                    torch.cuda.synchronize()  # before the next update rewrites them
            self.grad_evals = total
        else:
            ge = C.c_int64()
            _lib.check(self.lib.nax_run(self.h, int(target), C.byref(col), C.byref(ge)),
                       "nax_run")
            self.grad_evals = ge.value
        if "diverging" in out:
            out["diverging"] = out["diverging"].bool()
        return out

    def set_graphs(self, enable):
        """Replay nax_run's tick batches as hipGraphs (default) or launch
        every kernel directly; the draws are identical."""
        _lib.check(self.lib.nax_set_graphs(self.h, int(bool(enable))), "nax_set_graphs")

    def set_progress(self, fn, every_ticks=2048):
        """fn(ticks, pending) is called from the run loop every `every_ticks`
        gradient evaluations; None switches it off."""
        if fn is None:
            self._progress = None
            _lib.check(self.lib.nax_set_progress(self.h, _lib.PROGRESS_CB(), None, 0), "progress")
            return
        self._progress = _lib.PROGRESS_CB(lambda user, ticks, pending: fn(ticks, pending))
        _lib.check(self.lib.nax_set_progress(self.h, self._progress, None, int(every_ticks)),
                   "progress")

    def set_launch_shape(self, shape):
        """Fixes the launch shape the autotuner would pick (waves per chain of
        per-chain persistent models, chains per workgroup of the
        hierarchical model)."""
        _lib.check(self.lib.nax_set_launch_shape(self.h, int(shape)), "nax_set_launch_shape")

    def tuning(self):
        """The launch shapes the model's autotuner chose, as text."""
        buf = C.create_string_buffer(1024)
        _lib.check(self.lib.nax_tuning(self.h, buf, len(buf)), "nax_tuning")
        return buf.value.decode()

    def timing(self, enable):
        ms, n = C.c_double(), C.c_int64()
        _lib.check(self.lib.nax_kernel_timing(self.h, int(enable), C.byref(ms), C.byref(n)),
                   "timing")
        return ms.value, n.value

    # -- state --------------------------------------------------------------------
    def get_state(self):
        Cn, D = self.C, self.D
        t = {k: torch.empty((Cn, D), dtype=torch.float32, device="cuda")
             for k in ("z", "z_grad", "imm", "msqrt")}
        if self.dense:
            for k in ("imm", "msqrt"):
                t[k] = torch.empty((Cn, D, D), dtype=torch.float32, device="cuda")
        h = {"i": np.zeros(Cn, np.int32), "pe": np.zeros(Cn), "energy": np.zeros(Cn),
             "ap": np.zeros(Cn), "map": np.zeros(Cn), "ns": np.zeros(Cn, np.int32),
             "div": np.zeros(Cn, np.int32), "ss": np.zeros(Cn),
             "key": np.zeros(2 * Cn, np.uint32)}
        v = _lib.nax_state_view(
            t["z"].data_ptr(), t["z_grad"].data_ptr(), t["imm"].data_ptr(),
            t["msqrt"].data_ptr(), h["i"].ctypes.data, h["pe"].ctypes.data,
            h["energy"].ctypes.data, h["ap"].ctypes.data, h["map"].ctypes.data,
            h["ns"].ctypes.data, h["div"].ctypes.data, h["ss"].ctypes.data,
            h["key"].ctypes.data)
        torch.cuda.synchronize()
        _lib.check(self.lib.nax_get_state(self.h, C.byref(v)), "nax_get_state")
        # mass_matrix_sqrt_inv: 1 / sqrt(M) (diagonal) or tril_inv = S^-1 (dense;
        # S is lower triangular: triangular solves, 64 chains at a time)
        if self.dense:
            msi = torch.empty_like(t["msqrt"])
            eye = torch.eye(D, dtype=torch.float64, device="cuda")
            for c0 in range(0, Cn, 64):
                blk = t["msqrt"][c0:c0 + 64].double()
                msi[c0:c0 + 64] = torch.linalg.solve_triangular(
                    blk, eye.expand_as(blk), upper=False).float()
        else:
            msi = 1.0 / t["msqrt"]
        ss_state, mm_state, window_idx, akey = self._get_adapt()
        adapt = HMCAdaptState(torch.from_numpy(h["ss"]), t["imm"], t["msqrt"],
                              msi, ss_state, mm_state, window_idx, akey)
        return HMCState(torch.from_numpy(h["i"]), t["z"], t["z_grad"],
                        torch.from_numpy(h["pe"]), torch.from_numpy(h["energy"]), None,
                        self.cfg.trajectory_length or None, torch.from_numpy(h["ns"]),
                        torch.from_numpy(h["ap"]), torch.from_numpy(h["map"]),
                        torch.from_numpy(h["div"]).bool(), adapt,
                        h["key"].reshape(Cn, 2))

    def _get_adapt(self):
        """(ss_state, mm_state, window_idx, rng_key) of HMCAdaptState
        (hmc_util.py:16-30): ss_state = (x_t, x_avg, g_avg, t, prox_center)
        [C] each, mm_state = (mean [C, D], m2 [C, D] or [C, D, D], n [C])."""
        Cn, D = self.C, self.D
        f = {k: np.zeros(Cn) for k in ("x", "xa", "ga", "pc")}
        ints = {k: np.zeros(Cn, np.int32) for k in ("t", "n", "w")}
        key = np.zeros(2 * Cn, np.uint32)
        mean = torch.empty((Cn, D), dtype=torch.float64, device="cuda")
        m2 = torch.empty((Cn, D, D) if self.dense else (Cn, D), dtype=torch.float64,
                         device="cuda")
        v = _lib.nax_adapt_view(f["x"].ctypes.data, f["xa"].ctypes.data, f["ga"].ctypes.data,
                                ints["t"].ctypes.data, f["pc"].ctypes.data, mean.data_ptr(),
                                m2.data_ptr(), ints["n"].ctypes.data, ints["w"].ctypes.data,
                                key.ctypes.data)
        torch.cuda.synchronize()
        _lib.check(self.lib.nax_get_adapt_state(self.h, C.byref(v)), "nax_get_adapt_state")
        t = torch.from_numpy
        ss = (t(f["x"]), t(f["xa"]), t(f["ga"]), t(ints["t"]), t(f["pc"]))
        return ss, (mean, m2, t(ints["n"])), t(ints["w"]), key.reshape(Cn, 2)

    def _set_adapt(self, adapt):
        Cn, D = self.C, self.D
        arr = lambda x, dt: np.ascontiguousarray(np.broadcast_to(  # noqa: E731
            np.asarray(torch.as_tensor(x).cpu().numpy(), dt), (Cn,)))
        ss, mm, widx, akey = adapt.ss_state, adapt.mm_state, adapt.window_idx, adapt.rng_key
        x, xa, ga, tt, pc = (arr(ss[0], np.float64), arr(ss[1], np.float64),
                             arr(ss[2], np.float64), arr(ss[3], np.int32), arr(ss[4], np.float64))
        shape2 = (Cn, D, D) if self.dense else (Cn, D)
        mean = torch.as_tensor(mm[0], dtype=torch.float64).to("cuda").expand(Cn, D).contiguous()
        m2 = torch.as_tensor(mm[1], dtype=torch.float64).to("cuda").expand(shape2).contiguous()
        n, w = arr(mm[2], np.int32), arr(widx, np.int32)
        key = np.ascontiguousarray(np.broadcast_to(
            np.asarray(akey, np.uint32).reshape(-1, 2), (Cn, 2))).ravel()
        v = _lib.nax_adapt_view(x.ctypes.data, xa.ctypes.data, ga.ctypes.data, tt.ctypes.data,
                                pc.ctypes.data, mean.data_ptr(), m2.data_ptr(), n.ctypes.data,
                                w.ctypes.data, key.ctypes.data)
        torch.cuda.synchronize()
        _lib.check(self.lib.nax_set_adapt_state(self.h, C.byref(v)), "nax_set_adapt_state")

    def set_state(self, st, rng_key=None):
        Cn = self.C
        z = torch.as_tensor(st.z, dtype=torch.float32).to("cuda").reshape(Cn, self.D).contiguous()
        zg = torch.as_tensor(st.z_grad, dtype=torch.float32).to("cuda").reshape(Cn, self.D).contiguous()
        imm = torch.as_tensor(st.adapt_state.inverse_mass_matrix, dtype=torch.float32).to("cuda")
        msq = torch.as_tensor(st.adapt_state.mass_matrix_sqrt, dtype=torch.float32).to("cuda")
        shape = (Cn, self.D, self.D) if self.dense else (Cn, self.D)
        imm = imm.expand(shape).contiguous()
        msq = msq.expand(shape).contiguous()
        keys = np.asarray(st.rng_key if rng_key is None else rng_key, np.uint32)
        keys = np.ascontiguousarray(np.broadcast_to(keys.reshape(-1, 2), (Cn, 2))).ravel()
        arr = lambda x, dt: np.ascontiguousarray(np.broadcast_to(  # noqa: E731
            np.asarray(torch.as_tensor(x).cpu().numpy(), dt), (Cn,)))
        i = arr(st.i, np.int32)
        pe = arr(st.potential_energy, np.float64)
        en = arr(st.energy, np.float64)
        ap = arr(st.accept_prob, np.float64)
        mp = arr(st.mean_accept_prob, np.float64)
        ns = arr(st.num_steps, np.int32)
        dv = arr(st.diverging, np.int32)
        ss = arr(st.adapt_state.step_size, np.float64)
        v = _lib.nax_state_view(z.data_ptr(), zg.data_ptr(), imm.data_ptr(), msq.data_ptr(),
                                i.ctypes.data, pe.ctypes.data, en.ctypes.data, ap.ctypes.data,
                                mp.ctypes.data, ns.ctypes.data, dv.ctypes.data, ss.ctypes.data,
                                keys.ctypes.data)
        torch.cuda.synchronize()
        _lib.check(self.lib.nax_set_state(self.h, C.byref(v)), "nax_set_state")
        if getattr(st.adapt_state, "ss_state", None) is not None:
            self._set_adapt(st.adapt_state)


def _as_model(potential_fn, init_params):
    if isinstance(potential_fn, Model):
        return potential_fn
    if callable(potential_fn):
        if init_params is None:
            raise ValueError("Valid value of `init_params` must be provided with `potential_fn`.")
        z = torch.as_tensor(init_params)
        return PotentialFn(potential_fn, z.shape[-1] if z.ndim > 0 else 1)
    raise ValueError("potential_fn must be a numpyro_amd model or a torch callable")


def hmc(potential_fn=None, potential_fn_gen=None, kinetic_fn=None, algo="NUTS"):
    """``numpyro/infer/hmc.py:113``: returns ``(init_kernel, sample_kernel)``.

    ``potential_fn`` is a built-in model (``numpyro_amd.models``) or a torch
    callable on one flat latent vector.  ``init_params`` carry a leading chain
    dimension when several chains are initialised at once."""
    if kinetic_fn is not None:
        raise ValueError("custom kinetic_fn is not supported (Euclidean kinetic energy only)")
    if algo not in {"HMC", "NUTS"}:
        raise ValueError("`algo` must be one of `HMC` or `NUTS`.")
    if potential_fn is not None and potential_fn_gen is not None:
        raise ValueError("Only one of `potential_fn` or `potential_fn_gen` must be provided.")
    box = {}

    def init_kernel(init_params, num_warmup, step_size=1.0, inverse_mass_matrix=None,
                    adapt_step_size=True, adapt_mass_matrix=True, dense_mass=False,
                    target_accept_prob=0.8, *, num_steps=None, trajectory_length=2 * math.pi,
                    max_tree_depth=10, find_heuristic_step_size=False,
                    forward_mode_differentiation=False, regularize_mass_matrix=True,
                    model_args=(), model_kwargs=None, rng_key=None):
        pf = potential_fn
        if potential_fn_gen is not None:
            pf = potential_fn_gen(*model_args, **(model_kwargs or {}))
        z0 = torch.as_tensor(np.asarray(init_params) if not torch.is_tensor(init_params)
                             else init_params, dtype=torch.float32)
        if z0.ndim == 0:
            z0 = z0.reshape(1)
        model = _as_model(pf, z0)
        if z0.ndim == 1:
            z0 = z0[None]
        num_chains = z0.shape[0]
        if rng_key is None:
            rng_key = nrandom.PRNGKey(0)
        keys = np.asarray(rng_key, dtype=np.uint32).reshape(-1, 2)
        if keys.shape[0] != num_chains:
            if keys.shape[0] == 1:
                keys = np.broadcast_to(keys, (num_chains, 2))
            else:
                raise ValueError("rng_key must hold one key per chain")
        eng = Engine(model, num_chains, algo, num_warmup, step_size=step_size,
                     adapt_step_size=adapt_step_size, adapt_mass_matrix=adapt_mass_matrix,
                     dense_mass=dense_mass, target_accept_prob=target_accept_prob,
                     num_steps=num_steps,
                     trajectory_length=trajectory_length if algo == "HMC" else None,
                     max_tree_depth=max_tree_depth,
                     find_heuristic_step_size=find_heuristic_step_size,
                     regularize_mass_matrix=regularize_mass_matrix)
        eng.init(z0, keys, inverse_mass_matrix)
        box["engine"] = eng
        st = eng.get_state()
        box["i"] = int(st.i.max())
        return st

    def sample_kernel(hmc_state, model_args=(), model_kwargs=None):
        eng = box["engine"]
        i = int(torch.as_tensor(hmc_state.i).max())
        if i != box["i"]:
            eng.set_state(hmc_state)
        eng.run(i + 1)
        st = eng.get_state()
        box["i"] = i + 1
        return st

    init_kernel.engine = lambda: box.get("engine")
    sample_kernel.engine = lambda: box.get("engine")
    sample_kernel._nax = True
    return init_kernel, sample_kernel


class HMC(MCMCKernel):
    """``numpyro/infer/hmc.py:541-822``."""

    def __init__(self, model=None, potential_fn=None, kinetic_fn=None, step_size=1.0,
                 inverse_mass_matrix=None, adapt_step_size=True, adapt_mass_matrix=True,
                 dense_mass=False, target_accept_prob=0.8, num_steps=None,
                 trajectory_length=2 * math.pi, init_strategy=None,
                 find_heuristic_step_size=False, forward_mode_differentiation=False,
                 regularize_mass_matrix=True):
        if not (model is None) ^ (potential_fn is None):
            raise ValueError("Only one of `model` or `potential_fn` must be specified.")
        if type(self) is HMC:
            if num_steps is None and trajectory_length is None:
                raise ValueError(
                    "At least one of `num_steps` or `trajectory_length` must be specified.")
            if adapt_step_size and num_steps is not None and trajectory_length is not None:
                warnings.warn("If both `num_steps` and `trajectory_length` are specified step "
                              "size can't be adapted", stacklevel=2)
        if kinetic_fn is not None:
            raise ValueError("custom kinetic_fn is not supported (Euclidean kinetic energy only)")
        self._model = model
        self._potential_fn = potential_fn
        self._num_steps = num_steps
        self._step_size = float(step_size)
        self._inverse_mass_matrix = inverse_mass_matrix
        self._adapt_step_size = adapt_step_size
        self._adapt_mass_matrix = adapt_mass_matrix
        self._dense_mass = dense_mass
        self._target_accept_prob = target_accept_prob
        self._trajectory_length = trajectory_length
        self._algo = "HMC"
        self._max_tree_depth = 10
        self._find_heuristic_step_size = find_heuristic_step_size
        self._regularize_mass_matrix = regularize_mass_matrix
        self._init_strategy = init_strategy
        self._engine = None
        self._postprocess_fn = None
        self._deterministic_sites = ()

    @property
    def model(self):
        return self._model

    @property
    def sample_field(self):
        return "z"

    @property
    def default_fields(self):
        return ("z", "diverging")

    def get_diagnostics_str(self, state):
        return "{} steps of size {:.2e}. acc. prob={:.2f}".format(
            int(torch.as_tensor(state.num_steps).max()),
            float(torch.as_tensor(state.adapt_state.step_size).mean()),
            float(torch.as_tensor(state.mean_accept_prob).mean()))

    def _target_model(self):
        return self._model if self._model is not None else self._potential_fn

    def init(self, rng_key, num_warmup, init_params=None, model_args=(), model_kwargs={}):
        """``hmc.py:740-800``.  ``rng_key`` is one key or [C, 2] chain keys."""
        target = self._target_model()
        keys = np.asarray(rng_key, dtype=np.uint32).reshape(-1, 2)
        num_chains = keys.shape[0]
        if self._model is not None and is_model_function(self._model):
            # NUTS(model): trace the model once and bind the matching kernels
            # (the reference's initialize_model, numpyro/infer/util.py:663-835)
            target = compile_model(self._model, *model_args, **(model_kwargs or {}))
            if target.params.get("subsample"):
                # numpyro/primitives.py:448-456
                raise ValueError(
                    "Missing random key to generate subsample indices. Algorithms like HMC/NUTS "
                    "do not support subsampling. You might want to use SVI or HMCECS instead.")
            target.bind(num_chains)
            self._deterministic_sites = tuple(target.deterministic)
            if isinstance(init_params, dict):
                init_params = target.flatten(init_params)
        user_params = init_params is not None
        init_values = None
        if init_params is None:
            if not isinstance(target, Model):
                raise ValueError(
                    "Valid value of `init_params` must be provided with `potential_fn`.")
            if self._init_strategy is not None:
                if not hasattr(self._init_strategy, "unconstrained"):
                    raise ValueError("init_strategy must be init_to_value(values=...) or "
                                     "init_to_feasible() from numpyro_amd.infer")
                init_values = self._init_strategy.unconstrained(target.site_sizes)
            init_params = torch.zeros((num_chains, target.dim))  # drawn below
        z0 = torch.as_tensor(init_params, dtype=torch.float32)
        if z0.ndim <= 1 and num_chains == 1:
            z0 = z0.reshape(1, -1)
        model = _as_model(target, z0[0] if z0.ndim > 1 else z0)
        init_keys = np.stack([nrandom.split(k)[0] for k in keys]) if num_chains <= 8 else \
            _split_first(keys)
        eng = Engine(model, num_chains, self._algo, num_warmup, step_size=self._step_size,
                     adapt_step_size=self._adapt_step_size,
                     adapt_mass_matrix=self._adapt_mass_matrix, dense_mass=self._dense_mass,
                     target_accept_prob=self._target_accept_prob, num_steps=self._num_steps,
                     trajectory_length=self._trajectory_length if self._algo == "HMC" else None,
                     max_tree_depth=self._max_tree_depth,
                     find_heuristic_step_size=self._find_heuristic_step_size,
                     regularize_mass_matrix=self._regularize_mass_matrix)
        z0 = z0.reshape(num_chains, -1)
        if not getattr(model, "valid_obs", True):
            raise RuntimeError("Cannot find valid initial parameters. Please check your model "
                               "again.")
        if not user_params and not isinstance(model, PotentialFn):
            if init_values == "feasible":  # init_to_feasible: every coordinate 0
                z0 = torch.zeros((num_chains, model.dim), dtype=torch.float32)
            else:
                z0 = _initial_params(model, keys, values=init_values, eng=eng)
        eng.init(z0, init_keys, self._inverse_mass_matrix)
        self._engine = eng
        if isinstance(model, Model) and not isinstance(model, PotentialFn):
            self._postprocess_fn = model.postprocess
        return eng.get_state()

    def postprocess_fn(self, args, kwargs):
        if self._postprocess_fn is None:
            return lambda x: x
        return self._postprocess_fn

    def sample(self, state, model_args, model_kwargs):
        """One iteration of every chain (``hmc.py:806-816``)."""
        eng = self._engine
        i = int(torch.as_tensor(state.i).max())
        cur = eng.get_state()
        if int(cur.i.max()) != i:
            eng.set_state(state)
        eng.run(i + 1)
        return eng.get_state()

    def __getstate__(self):
        st = self.__dict__.copy()
        st["_engine"] = None
        return st


class NUTS(HMC):
    """``numpyro/infer/hmc.py:825-951``."""

    def __init__(self, model=None, potential_fn=None, kinetic_fn=None, step_size=1.0,
                 inverse_mass_matrix=None, adapt_step_size=True, adapt_mass_matrix=True,
                 dense_mass=False, target_accept_prob=0.8, trajectory_length=None,
                 max_tree_depth=10, init_strategy=None, find_heuristic_step_size=False,
                 forward_mode_differentiation=False, regularize_mass_matrix=True):
        super().__init__(model=model, potential_fn=potential_fn, kinetic_fn=kinetic_fn,
                         step_size=step_size, inverse_mass_matrix=inverse_mass_matrix,
                         adapt_step_size=adapt_step_size, adapt_mass_matrix=adapt_mass_matrix,
                         dense_mass=dense_mass, target_accept_prob=target_accept_prob,
                         trajectory_length=trajectory_length,
                         find_heuristic_step_size=find_heuristic_step_size,
                         regularize_mass_matrix=regularize_mass_matrix,
                         init_strategy=init_strategy)
        self._max_tree_depth = max_tree_depth
        self._algo = "NUTS"


def _split_first(keys):
    """split(k)[0] for each key, on the device."""
    out = np.empty_like(keys)
    for c in range(keys.shape[0]):
        out[c] = nrandom.split(keys[c])[0]
    return out


def _site_layout(model):
    """The latent sites in the model's execution order, each with its
    offset and size in the flat (sorted-name) latent and its transform."""
    sizes = tuple(getattr(model, "site_sizes", ()) or ())
    if not sizes:  # one unnamed block (ragged models): the whole latent
        return [("_latent", 0, int(model.dim), None)]
    offs, o = {}, 0
    for name, size, tf in sizes:
        offs[name] = (o, int(size), tf)
        o += int(size)
    order = tuple(getattr(model, "site_order", ()) or ()) or tuple(n for n, _, _ in sizes)
    return [(name,) + offs[name] for name in order]


class _InitKeys:
    """find_valid_initial_params for init_to_uniform / init_to_value
    (numpyro/infer/util.py:405-466, its prototype branch): per chain, the
    key handed to initialize_model is the second half of split(chain key)
    (HMC.init, hmc.py:744-750); every attempt draws ``key, subkey =
    split(key)`` and then, site by site in the model's execution order,
    ``uniform(subkey, shape, -radius, radius)`` followed by ``key, subkey =
    split(key)``.  Sites given by init_to_value take their (unconstrained)
    values and consume no key.  The next attempt continues from the key
    the previous one ended with."""

    def __init__(self, model, keys, radius=2.0, values=None):
        self.layout = _site_layout(model)
        self.dim = int(model.dim)
        self.radius = np.float32(radius)
        self.values = dict(values or {})
        keys = np.asarray(keys, dtype=np.uint32).reshape(-1, 2)
        self.keys = _split_at(keys, 1)

    def draw(self, c):
        lo = -self.radius
        key = self.keys[c]
        key, sub = nrandom.split(key)
        z = np.zeros(self.dim, np.float32)
        for name, off, size, _ in self.layout:
            if name in self.values:
                v = self.values[name]
                z[off:off + size] = v[c if v.shape[0] > 1 else 0]
                continue
            u = _uniform(sub, size).cpu().numpy()
            # float32 f * (max - min) + min, clamped below by min
            z[off:off + size] = np.maximum(lo, u * (self.radius - lo) + lo)
            key, sub = nrandom.split(key)
        self.keys[c] = key
        return z


def _split_at(keys, i):
    """split(k)[i] for each key (one batched device call)."""
    kd = torch.as_tensor(np.ascontiguousarray(keys).view(np.int32)).cuda().contiguous()
    out = torch.empty((keys.shape[0], 2, 2), dtype=torch.int32, device="cuda")
    _lib.check(_lib.load().nax_split_keys(_lib.ptr(kd), keys.shape[0], 2,
                                          int(nrandom.PARTITIONABLE), _lib.ptr(out)),
               "nax_split_keys")
    return out[:, i].cpu().numpy().view(np.uint32).copy()


def _initial_params(model, keys, values=None, radius=2.0, max_tries=100, eng=None):
    """Initial points of all chains as find_valid_initial_params draws them,
    redrawn (continuing each chain's key sequence) while U or dU/dz is not
    finite there, at most 100 attempts (numpyro/infer/util.py:417-466)."""
    draws = _InitKeys(model, keys, radius, values)
    C_ = np.asarray(keys).reshape(-1, 2).shape[0]
    z = torch.from_numpy(np.stack([draws.draw(c) for c in range(C_)]))
    if eng is None:
        return z
    D = z.shape[1]
    ld = _lib.stride(eng.h)
    lib = _lib.load()
    for _ in range(max_tries):
        zp = torch.zeros((C_, ld), dtype=torch.float32, device="cuda")
        zp[:, :D] = z.to("cuda")
        u = torch.empty(C_, dtype=torch.float64, device="cuda")
        g = torch.zeros((C_, ld), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        _lib.check(lib.nax_potential_grad(eng.h, _lib.ptr(zp), ld, C_, _lib.ptr(u), _lib.ptr(g)),
                   "nax_potential_grad")
        bad = (~(torch.isfinite(u) & torch.isfinite(g[:, :D]).all(dim=1))).cpu()
        if not bool(bad.any()):
            return z
        for c in torch.nonzero(bad).flatten().tolist():
            z[c] = torch.from_numpy(draws.draw(c))
    raise RuntimeError("Cannot find valid initial parameters. Please check your model again.")


def init_to_uniform(keys, dim, radius=2.0, self_split=False):
    """Uniform(-radius, radius) draws of a flat latent of size ``dim``, one
    per chain key: ``uniform(subkey, (dim,))`` with ``subkey =
    split(k)[1]`` (``self_split=True`` first takes the second half of
    ``split(key)`` as HMC.init does, hmc.py:744-750).  For models,
    ``HMC.init`` draws site by site as find_valid_initial_params does."""
    keys = np.asarray(keys, dtype=np.uint32).reshape(-1, 2)
    rows = []
    for k in keys:
        kk = nrandom.split(k)[1] if self_split else k
        sub = nrandom.split(kk)[1]
        u = _uniform(sub, dim)
        rows.append(torch.clamp(u * (2 * radius) - radius, min=-radius))
    return torch.stack(rows)


def _uniform(k, n):
    # uniform [0,1) float32 from the device normal path's bits: u = (x + 1) / 2
    # is not bit-identical to random.uniform; derive from threefry bits instead.
    lib = _lib.load()
    dk = nrandom._dev_u32(k)
    x0 = torch.zeros(n, dtype=torch.int32, device="cuda")
    x1 = torch.arange(n, dtype=torch.int32, device="cuda")
    o0 = torch.empty_like(x0)
    o1 = torch.empty_like(x0)
    _lib.check(lib.nax_threefry2x32(_lib.ptr(dk), _lib.ptr(x0), _lib.ptr(x1), _lib.ptr(o0),
                                    _lib.ptr(o1), n), "threefry")
    bits = (o0 ^ o1)
    b = ((bits >> 9) & 0x7FFFFF) | 0x3F800000
    return b.view(torch.float32) - 1.0
