"""ctypes wrapper of oracle/libcpu_nuts.so -- test / baseline only."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libcpu_nuts.so")


def _lib():
    lib = C.CDLL(LIB)
    f = lib.cpu_nuts_logistic
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_double, C.c_int, C.c_void_p,
                  C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_void_p,
                  C.c_void_p, C.POINTER(C.c_int64), C.c_int]
    return lib


def nuts_logistic(X, y, z0, keys, num_warmup, num_samples, max_depth=10, step_size=1.0,
                  adapt=True, prior_scale=1.0, threads=0):
    X = np.ascontiguousarray(X, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    z0 = np.ascontiguousarray(z0, dtype=np.float32)
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    Cn, D = z0.shape
    z_out = np.zeros((Cn, num_samples, D), np.float32)
    ns = np.zeros((Cn, num_warmup + num_samples), np.int32)
    ge = C.c_int64()
    rc = _lib().cpu_nuts_logistic(X.ctypes.data, y.ctypes.data, X.shape[0], D, prior_scale, Cn,
                                  z0.ctypes.data, keys.ctypes.data, num_warmup, num_samples,
                                  max_depth, step_size, int(adapt), z_out.ctypes.data,
                                  ns.ctypes.data, C.byref(ge), threads)
    assert rc == 0
    return z_out, ns, ge.value


def logistic_batch_grad(X, y, z, prior_scale=1.0, threads=0):
    """U [C] and dU/dz [C, D] (float64) of the logistic model for the chains
    at z [C, D], one multi-threaded pass over X for all chains."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    z = np.ascontiguousarray(z, dtype=np.float32)
    Cn, D = z.shape
    U = np.zeros(Cn)
    g = np.zeros((Cn, D))
    lib = C.CDLL(LIB)
    f = lib.cpu_logistic_batch_grad
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_double, C.c_int, C.c_void_p,
                  C.c_void_p, C.c_void_p, C.c_int]
    rc = f(X.ctypes.data, y.ctypes.data, X.shape[0], D, prior_scale, Cn, z.ctypes.data,
           U.ctypes.data, g.ctypes.data, threads)
    assert rc == 0
    return U, g
