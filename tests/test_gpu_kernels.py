"""GPU parity of the engine's primitives against the oracle: threefry bits
(bit-exact), key derivations (bit-exact), normals (float tolerance), the
logistic potential + adjoint (float64 oracle) and L-step leapfrog
trajectories (float64 oracle, 1e-5 relative)."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import hmc as ohmc
from oracle import hmc_util as ohu
from oracle import jax_random as ojr
from oracle import models as om

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from numpyro_amd import _lib
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return _lib


def _u32(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.uint32)).view(np.int32)).cuda()


def _np_u32(t):
    return t.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("key, ctr, expected", [
    ((0, 0), (0, 0), (0x6B200159, 0x99BA4EFE)),
    ((0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF), (0x1CB996FC, 0xBB002BE7)),
    ((0x13198A2E, 0x03707344), (0x243F6A88, 0x85A308D3), (0xC4923A9C, 0x483DF7A0)),
])
def test_threefry_kat_gpu(lib, key, ctr, expected):
    k, x0, x1 = _u32(key), _u32([ctr[0]]), _u32([ctr[1]])
    o0, o1 = torch.empty_like(x0), torch.empty_like(x0)
    lib.check(lib.load().nax_threefry2x32(lib.ptr(k), lib.ptr(x0), lib.ptr(x1), lib.ptr(o0),
                                          lib.ptr(o1), 1))
    assert (int(_np_u32(o0)[0]), int(_np_u32(o1)[0])) == expected


def test_threefry_bulk_matches_oracle(lib):
    rng = np.random.default_rng(0)
    n = 100_003
    key = rng.integers(0, 2**32, 2, dtype=np.uint64).astype(np.uint32)
    a = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    k, x0, x1 = _u32(key), _u32(a), _u32(b)
    o0, o1 = torch.empty_like(x0), torch.empty_like(x0)
    lib.check(lib.load().nax_threefry2x32(lib.ptr(k), lib.ptr(x0), lib.ptr(x1), lib.ptr(o0),
                                          lib.ptr(o1), n))
    e0, e1 = ojr.threefry2x32(key, a, b)
    assert np.array_equal(_np_u32(o0), e0) and np.array_equal(_np_u32(o1), e1)


@pytest.mark.parametrize("part", [True, False])
@pytest.mark.parametrize("num", [1, 2, 3, 7, 1024])
def test_split_matches_oracle(lib, part, num):
    from numpyro_amd import random as nr
    key = np.array([123, 456], np.uint32)
    got = nr.split(key, num, partitionable=part)
    assert np.array_equal(got, ojr.split(key, num, partitionable=part))


@pytest.mark.parametrize("part", [True, False])
def test_chain_keys_match_oracle(lib, part):
    from numpyro_amd import random as nr
    for C in (1, 4, 33):
        key = ojr.PRNGKey(7)
        got = nr.chain_init_keys(key, C, partitionable=part)
        old = ojr.PARTITIONABLE
        ojr.PARTITIONABLE = part
        try:
            want = ohmc.chain_keys(key, C)
        finally:
            ojr.PARTITIONABLE = old
        assert np.array_equal(got, want)


@pytest.mark.parametrize("part", [True, False])
@pytest.mark.parametrize("n", [1, 26, 255, 4096])
def test_normal_matches_oracle(lib, part, n):
    from numpyro_amd import random as nr
    key = np.array([9, 10], np.uint32)
    got = nr.normal(key, n, partitionable=part).cpu().numpy().astype(np.float64)
    want = ojr.normal(key, n, partitionable=part)
    # float32 erfinv (Giles' single-precision approximation, relative error
    # up to ~6e-6 in its tail branch) vs the float64 value at the same
    # float32 uniform
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=2e-6)


def _logistic_problem(N, D, C, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(N, D)).astype(np.float32)
    beta = rng.normal(size=D) * 0.5
    y = (rng.uniform(size=N) < 1 / (1 + np.exp(-X @ beta))).astype(np.float32)
    Z = (beta[None] + 0.1 * rng.normal(size=(C, D))).astype(np.float32)
    return X, y, Z


def _engine_for(X, y, C, x_dtype="fp32", prior_scale=1.0):
    from numpyro_amd.infer.hmc import Engine
    from numpyro_amd.models import LogisticRegression
    m = LogisticRegression(X, y, prior_scale=prior_scale, x_dtype=x_dtype)
    return Engine(m, C, "NUTS", 0)


@pytest.mark.parametrize("N, D, C", [(1000, 26, 4), (1, 3, 1), (777, 55, 9), (20000, 55, 64),
                                     (3001, 256, 16), (100, 1, 3)])
def test_logistic_potential_grad_fp32(lib, N, D, C):
    X, y, Z = _logistic_problem(N, D, C, N + D + C)
    eng = _engine_for(X, y, C, prior_scale=1.3)
    z = torch.from_numpy(Z).cuda()
    pe = torch.empty(C, dtype=torch.float64, device="cuda")
    g = torch.empty((C, D), dtype=torch.float32, device="cuda")
    lib.check(lib.load().nax_potential_grad(eng.h, lib.ptr(z), D, C, lib.ptr(pe), lib.ptr(g)))
    for c in range(C):
        u, gr = om.logistic_potential_grad(Z[c].astype(np.float64), X, y, 1.3)
        np.testing.assert_allclose(pe[c].item(), u, rtol=1e-5)
        scale = np.abs(gr).max() + 1.0
        np.testing.assert_allclose(g[c].cpu().numpy(), gr, rtol=1e-4, atol=1e-5 * scale)


@pytest.mark.parametrize("N, D, C, waves", [(1000, 26, 4, None), (1, 3, 1, None),
                                            (777, 55, 9, None), (100, 1, 3, None),
                                            (5000, 64, 2, None), (63, 33, 5, None),
                                            (1000, 26, 4, 4), (777, 55, 9, 2), (5000, 64, 2, 16),
                                            (63, 33, 5, 8), (1, 3, 1, 16)])
def test_logistic_potential_grad_chain_kernel(lib, N, D, C, waves):
    # a group of waves per chain (kernel "chain"): rows split over the
    # group's threads, d <= 64; waves fixed by the launch shape, else one
    from numpyro_amd.infer.hmc import Engine
    from numpyro_amd.models import LogisticRegression
    X, y, Z = _logistic_problem(N, D, C, N + D + C + 1)
    eng = Engine(LogisticRegression(X, y, prior_scale=0.9, kernel="chain"), C, "NUTS", 0)
    if waves:
        eng.set_launch_shape(waves)
    z = torch.from_numpy(Z).cuda()
    pe = torch.empty(C, dtype=torch.float64, device="cuda")
    g = torch.empty((C, D), dtype=torch.float32, device="cuda")
    lib.check(lib.load().nax_potential_grad(eng.h, lib.ptr(z), D, C, lib.ptr(pe), lib.ptr(g)))
    for c in range(C):
        u, gr = om.logistic_potential_grad(Z[c].astype(np.float64), X, y, 0.9)
        np.testing.assert_allclose(pe[c].item(), u, rtol=1e-5)
        scale = np.abs(gr).max() + 1.0
        np.testing.assert_allclose(g[c].cpu().numpy(), gr, rtol=1e-4, atol=1e-5 * scale)


def test_leapfrog_trajectory_matches_oracle(lib):
    # L-step velocity_verlet from identical (q, p, eps, M): 1e-5 relative
    N, D, C, L = 1000, 26, 4, 25
    X, y, Z = _logistic_problem(N, D, C, 5)
    rng = np.random.default_rng(6)
    R = rng.normal(size=(C, D)).astype(np.float32)
    imm = (0.5 + rng.uniform(size=D)).astype(np.float32) * 0.05
    eps = 0.05
    eng = _engine_for(X, y, C)
    z = torch.from_numpy(Z.copy()).cuda()
    r = torch.from_numpy(R.copy()).cuda()
    g = torch.empty((C, D), dtype=torch.float32, device="cuda")
    pe = torch.empty(C, dtype=torch.float64, device="cuda")
    lib.check(lib.load().nax_potential_grad(eng.h, lib.ptr(z), D, C, lib.ptr(pe), lib.ptr(g)))
    m = torch.from_numpy(imm).cuda()
    lib.check(lib.load().nax_leapfrog(eng.h, lib.ptr(z), lib.ptr(r), lib.ptr(g), lib.ptr(pe),
                                      eps, L, lib.ptr(m)))
    f = lambda q: om.logistic_potential_grad(q, X, y, 1.0)  # noqa: E731
    init, update = ohu.velocity_verlet(f)
    for c in range(C):
        st = init(Z[c].astype(np.float64), R[c].astype(np.float64))
        for _ in range(L):
            st = update(eps, imm.astype(np.float64), st)
        zs = z[c].cpu().numpy()
        rs = r[c].cpu().numpy()
        np.testing.assert_allclose(zs, st.z, rtol=1e-5, atol=1e-5 * np.abs(st.z).max())
        np.testing.assert_allclose(rs, st.r, rtol=1e-5, atol=1e-5 * np.abs(st.r).max())
        np.testing.assert_allclose(pe[c].item(), st.potential_energy, rtol=1e-6)


def _bf16_round(a):
    """float32 -> bfloat16 (round to nearest even) -> float64."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def _eval(eng, lib, Z):
    C, D = Z.shape
    z = torch.from_numpy(Z).cuda()
    pe = torch.empty(C, dtype=torch.float64, device="cuda")
    g = torch.empty((C, D), dtype=torch.float32, device="cuda")
    lib.check(lib.load().nax_potential_grad(eng.h, lib.ptr(z), D, C, lib.ptr(pe), lib.ptr(g)))
    return pe.cpu().numpy(), g.cpu().numpy()


@pytest.mark.parametrize("N, D, C", [(3001, 256, 128), (1000, 55, 40), (777, 26, 33),
                                     (65, 32, 130), (4096, 200, 256), (50001, 256, 256),
                                     (20011, 55, 64),
                                     # one padded row tile, D = 1, one chain, ragged D
                                     (1, 1, 1), (5, 3, 2), (31, 33, 1)])
def test_logistic_mfma_fp32(lib, N, D, C):
    from numpyro_amd.infer.hmc import Engine
    from numpyro_amd.models import LogisticRegression
    X, y, Z = _logistic_problem(N, D, C, 7 * N + D)
    eng = Engine(LogisticRegression(X, y, prior_scale=0.7, kernel="mfma"), C, "NUTS", 0)
    pe, g = _eval(eng, lib, Z)
    for c in range(0, C, max(1, C // 9)):
        u, gr = om.logistic_potential_grad(Z[c].astype(np.float64), X, y, 0.7)
        np.testing.assert_allclose(pe[c], u, rtol=1e-5)
        scale = np.abs(gr).max() + 1.0
        np.testing.assert_allclose(g[c], gr, rtol=1e-4, atol=2e-5 * scale)


@pytest.mark.parametrize("N, D, C", [(3001, 256, 128), (1000, 55, 40), (777, 26, 33),
                                     (65, 32, 130), (4096, 200, 256),
                                     # many row tiles per block: exercises the LDS ring
                                     (50001, 256, 256), (20011, 55, 64), (9001, 26, 40),
                                     (70003, 160, 1024),
                                     # one padded row tile, D = 1, one chain, ragged D
                                     (1, 1, 1), (5, 3, 2), (31, 33, 1), (33, 255, 3)])
def test_logistic_mfma_bf16(lib, N, D, C):
    """Two-copy BF16 form (kernel 'two-copy').  BF16 storage defines the model
    with X rounded to bf16.  Against the
    float64 oracle on that X: U to 1e-5 relative (positions enter as a bf16
    hi+lo pair, ~2^-17 relative); dU within the rounding bound of the fp16
    R = sigmoid(l) - y of the second GEMM: |dg_d| <= 2^-11 sum_n |r_n x_nd|
    (unit roundoff of fp16) + 2^-25 sum_n |x_nd| (fp16 subnormals) + fp32
    accumulation.  The error against the unrounded X is the quantisation
    itself (reported by the bench)."""
    from numpyro_amd.infer.hmc import Engine
    from numpyro_amd.models import LogisticRegression
    from scipy.special import expit
    X, y, Z = _logistic_problem(N, D, C, 11 * N + D)
    eng = Engine(LogisticRegression(X, y, prior_scale=0.7, x_dtype="bf16", kernel="two-copy"), C,
                 "NUTS", 0)
    pe, g = _eval(eng, lib, Z)
    Xq = _bf16_round(X)
    for c in range(0, C, max(1, C // 9)):
        zc = Z[c].astype(np.float64)
        u, gr = om.logistic_potential_grad(zc, Xq, y, 0.7)
        np.testing.assert_allclose(pe[c], u, rtol=1e-5)
        r = expit(Xq @ zc) - y
        bound = (2.0 ** -11 * (np.abs(r)[:, None] * np.abs(Xq)).sum(0)
                 + 2.0 ** -25 * np.abs(Xq).sum(0) + 1e-5 * (np.abs(gr).max() + 1.0))
        err = np.abs(g[c] - gr)
        assert np.all(err <= bound), (err / bound).max()


@pytest.mark.parametrize("N, D, C", [(3001, 256, 128), (1000, 55, 40), (777, 26, 33),
                                     (65, 32, 130), (4096, 200, 256), (50001, 256, 256),
                                     (20011, 55, 64), (9001, 26, 40), (70003, 160, 1024),
                                     (1, 1, 1), (5, 3, 2), (31, 33, 1), (33, 255, 3)])
def test_logistic_mfma_bf16_single_copy(lib, N, D, C):
    """Single-copy bf16 form (kernel 'single', the bf16 default): X^T R reads the bf16 copy of
    X transposed (ds_read_b64_tr_b16) and R = sigmoid(l) - y is rounded to
    bf16.  U as the two-copy form (1e-5 relative on the bf16-rounded X); dU
    within |dg_d| <= 2^-8 sum_n |r_n x_nd| (bf16 unit roundoff of R) + fp32
    accumulation."""
    from numpyro_amd.infer.hmc import Engine
    from numpyro_amd.models import LogisticRegression
    from scipy.special import expit
    X, y, Z = _logistic_problem(N, D, C, 13 * N + D)
    eng = Engine(LogisticRegression(X, y, prior_scale=0.7, x_dtype="bf16", kernel="single"), C,
                 "NUTS", 0)
    pe, g = _eval(eng, lib, Z)
    Xq = _bf16_round(X)
    for c in range(0, C, max(1, C // 9)):
        zc = Z[c].astype(np.float64)
        u, gr = om.logistic_potential_grad(zc, Xq, y, 0.7)
        np.testing.assert_allclose(pe[c], u, rtol=1e-5)
        r = expit(Xq @ zc) - y
        bound = (2.0 ** -8 * (np.abs(r)[:, None] * np.abs(Xq)).sum(0)
                 + 1e-5 * (np.abs(gr).max() + 1.0))
        err = np.abs(g[c] - gr)
        assert np.all(err <= bound), (err / bound).max()


def test_single_copy_needs_bf16(lib):
    from numpyro_amd.models import LogisticRegression
    X, y, _ = _logistic_problem(10, 3, 1, 1)
    with pytest.raises(ValueError, match="x_dtype='bf16'"):
        LogisticRegression(X, y, kernel="single")
