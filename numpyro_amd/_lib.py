"""ctypes binding of libnumpyro_amd.so (the C ABI in include/numpyro_amd.h).

The library is loaded from this package directory.  There is no fallback:
if the library is missing, or no HIP device is present, calls raise.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnumpyro_amd.so")

NAX_ALGO_NUTS = 0
NAX_ALGO_HMC = 1
X_DTYPES = {"fp32": 0, "float32": 0, "f32": 0, "bf16": 1, "bfloat16": 1, "fp8": 2,
            "fp8_e4m3": 2}


class nax_config(C.Structure):
    _fields_ = [
        ("num_chains", C.c_int32), ("dim", C.c_int32), ("algo", C.c_int32),
        ("num_warmup", C.c_int32), ("max_tree_depth_warmup", C.c_int32),
        ("max_tree_depth", C.c_int32), ("adapt_step_size", C.c_int32),
        ("adapt_mass_matrix", C.c_int32), ("regularize_mass_matrix", C.c_int32),
        ("partitionable_keys", C.c_int32), ("num_steps", C.c_int32),
        ("find_heuristic_step_size", C.c_int32),
        ("step_size", C.c_double), ("target_accept_prob", C.c_double),
        ("trajectory_length", C.c_double), ("max_delta_energy", C.c_double),
        ("dense_mass", C.c_int32), ("reserved", C.c_int32),
    ]


class nax_collect(C.Structure):
    _fields_ = [
        ("start_idx", C.c_int64), ("thinning", C.c_int32), ("size", C.c_int32),
        ("d_z", C.c_void_p), ("d_potential_energy", C.c_void_p), ("d_energy", C.c_void_p),
        ("d_accept_prob", C.c_void_p), ("d_mean_accept_prob", C.c_void_p),
        ("d_num_steps", C.c_void_p), ("d_diverging", C.c_void_p), ("d_step_size", C.c_void_p),
    ]


class nax_state_view(C.Structure):
    _fields_ = [
        ("d_z", C.c_void_p), ("d_z_grad", C.c_void_p), ("d_inverse_mass_matrix", C.c_void_p),
        ("d_mass_matrix_sqrt", C.c_void_p), ("i", C.c_void_p), ("potential_energy", C.c_void_p),
        ("energy", C.c_void_p), ("accept_prob", C.c_void_p), ("mean_accept_prob", C.c_void_p),
        ("num_steps", C.c_void_p), ("diverging", C.c_void_p), ("step_size", C.c_void_p),
        ("rng_key", C.c_void_p),
    ]


class nax_adapt_view(C.Structure):
    _fields_ = [
        ("ss_x_t", C.c_void_p), ("ss_x_avg", C.c_void_p), ("ss_g_avg", C.c_void_p),
        ("ss_t", C.c_void_p), ("ss_prox_center", C.c_void_p), ("d_mm_mean", C.c_void_p),
        ("d_mm_m2", C.c_void_p), ("mm_n", C.c_void_p), ("window_idx", C.c_void_p),
        ("rng_key", C.c_void_p),
    ]


PROGRESS_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int64, C.c_int32)

# name -> (restype, argtypes)
SIGNATURES = {
    "nax_last_error": (C.c_char_p, []),
    "nax_create": (C.c_int, [C.POINTER(nax_config), C.POINTER(C.c_void_p)]),
    "nax_destroy": (None, [C.c_void_p]),
    "nax_model_logistic": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32,
                                     C.c_double, C.c_int32]),
    "nax_model_logistic_ex": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                        C.c_int32, C.c_double, C.c_int32, C.c_int32]),
    "nax_model_sv": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32]),
    "nax_model_sv_ex": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_double,
                                  C.c_double]),
    "nax_model_hier": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_int32, C.c_int32]),
    "nax_model_external": (C.c_int, [C.c_void_p]),
    "nax_set_positive_transform": (C.c_int, [C.c_void_p, C.c_int32]),
    "nax_tuning": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32]),
    "nax_set_launch_shape": (C.c_int, [C.c_void_p, C.c_int32]),
    "nax_fp8_quantize": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p]),
    "nax_model_logistic_ecs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                         C.c_int32, C.c_double, C.c_int32, C.c_int32, C.c_int32,
                                         C.c_void_p, C.c_void_p, C.c_void_p]),
    "nax_ecs_gibbs": (C.c_int, [C.c_void_p, C.c_void_p]),
    "nax_randint": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                              C.c_void_p]),
    "nax_split_keys": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]),
    "nax_init": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "nax_run": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(nax_collect),
                          C.POINTER(C.c_int64)]),
    "nax_advance": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(nax_collect),
                              C.POINTER(C.c_int32)]),
    "nax_buffers": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                              C.POINTER(C.c_void_p), C.POINTER(C.c_int32)]),
    "nax_exchange": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "nax_get_state": (C.c_int, [C.c_void_p, C.POINTER(nax_state_view)]),
    "nax_set_state": (C.c_int, [C.c_void_p, C.POINTER(nax_state_view)]),
    "nax_get_adapt_state": (C.c_int, [C.c_void_p, C.POINTER(nax_adapt_view)]),
    "nax_set_adapt_state": (C.c_int, [C.c_void_p, C.POINTER(nax_adapt_view)]),
    "nax_potential_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                     C.c_void_p, C.c_void_p]),
    "nax_logistic_potential_grad_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64,
                                                      C.c_int32, C.c_double, C.c_void_p,
                                                      C.c_int32, C.c_int32, C.c_void_p,
                                                      C.c_void_p, C.c_void_p]),
    "nax_leapfrog": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_double, C.c_int32, C.c_void_p]),
    "nax_threefry2x32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_int64]),
    "nax_random_split": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]),
    "nax_random_normal": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]),
    "nax_chain_keys": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]),
    "nax_set_progress": (C.c_int, [C.c_void_p, PROGRESS_CB, C.c_void_p, C.c_int64]),
    "nax_set_graphs": (C.c_int, [C.c_void_p, C.c_int32]),
    "nax_diagnostics": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                  C.c_void_p]),
    "nax_kernel_timing": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double),
                                    C.POINTER(C.c_int64)]),
}

_lib = None


def load():
    """Loads the library (without touching the GPU) and binds every symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m numpyro_amd.build` "
            "(numpyro_amd has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class NaxError(RuntimeError):
    pass


def check(rc, what=""):
    if rc != 0:
        msg = load().nax_last_error().decode(errors="replace")
        if rc == -22:
            raise ValueError(f"{what}: {msg}")
        raise NaxError(f"{what} failed ({rc}): {msg}")


def ptr(t):
    """Device/host pointer of a torch tensor (or None)."""
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def stride(handle):
    """Leading dimension (Dp) of the engine's per-chain vectors."""
    ld = C.c_int32()
    check(load().nax_buffers(handle, None, None, None, C.byref(ld)), "nax_buffers")
    return ld.value
