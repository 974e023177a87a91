"""The allocation-free entry point nax_logistic_potential_grad_noalloc against
nax_potential_grad on a sampler handle with the same kernels (the logistic
model's fp32 tile form): two calls in a row, the second on a smaller input,
each bit-identical to the handle-based result (same kernels, same row-split
order), and sizes beyond the scratch array refused."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _noalloc(X, y, z, ld, prior_scale=1.0):
    from numpyro_amd import _lib
    count = z.shape[0]
    pe = torch.full((count,), float("nan"), dtype=torch.float64, device="cuda")
    g = torch.full((count, ld), float("nan"), dtype=torch.float32, device="cuda")
    _lib.check(_lib.load().nax_logistic_potential_grad_noalloc(
        _lib.ptr(X), _lib.ptr(y), X.shape[0], X.shape[1], prior_scale, _lib.ptr(z), ld, count,
        _lib.ptr(pe), _lib.ptr(g), None), "noalloc")
    torch.cuda.synchronize()
    return pe, g


def _handle(X, y, z, ld, prior_scale=1.0):
    from numpyro_amd import _lib
    from numpyro_amd.infer.hmc import Engine
    from numpyro_amd.models import LogisticRegression
    count = z.shape[0]
    eng = Engine(LogisticRegression(X, y, prior_scale=prior_scale, kernel="tile"), count,
                 "NUTS", 0)
    pe = torch.empty(count, dtype=torch.float64, device="cuda")
    g = torch.full((count, ld), float("nan"), dtype=torch.float32, device="cuda")
    _lib.check(_lib.load().nax_potential_grad(eng.h, _lib.ptr(z), ld, count, _lib.ptr(pe),
                                              _lib.ptr(g)), "potential_grad")
    torch.cuda.synchronize()
    return pe, g


def _case(N, D, count, seed, prior_scale):
    rng = np.random.default_rng(seed)
    X = torch.from_numpy(rng.normal(size=(N, D)).astype(np.float32)).cuda()
    y = torch.from_numpy((rng.uniform(size=N) < 0.4).astype(np.float32)).cuda()
    ld = (D + 63) // 64 * 64
    z = torch.zeros((count, ld), dtype=torch.float32, device="cuda")
    z[:, :D] = torch.from_numpy(0.1 * rng.normal(size=(count, D)).astype(np.float32)).cuda()
    return X, y, z, ld


def test_noalloc_matches_the_handle_entry_point_on_two_calls():
    from numpyro_amd.models import LogisticRegression  # noqa: F401 (loads the library)
    for N, D, count, seed, s in [(20000, 55, 64, 0, 1.0), (3000, 7, 5, 1, 2.5)]:
        print(f"[noalloc] N={N} D={D} count={count}", flush=True)
        X, y, z, ld = _case(N, D, count, seed, s)
        pe_a, g_a = _noalloc(X, y, z, ld, s)
        pe_b, g_b = _handle(X, y, z, ld, s)
        assert torch.isfinite(pe_a).all() and torch.isfinite(g_a[:, :D]).all()
        assert torch.equal(pe_a, pe_b)
        assert torch.equal(g_a[:, :D], g_b[:, :D])


def test_noalloc_refuses_partials_beyond_the_scratch_array():
    from numpyro_amd import _lib
    lib = _lib.load()
    # 16384 chains x (8 + 4 * 320) bytes of partials: above 16 MiB
    x = torch.zeros((4, 320), dtype=torch.float32, device="cuda")
    y = torch.zeros(4, dtype=torch.float32, device="cuda")
    rc = lib.nax_logistic_potential_grad_noalloc(_lib.ptr(x), _lib.ptr(y), 4, 320, 1.0,
                                                 _lib.ptr(x), 320, 16384, _lib.ptr(y),
                                                 _lib.ptr(x), None)
    assert rc == -22
    assert b"16 MiB" in lib.nax_last_error()
