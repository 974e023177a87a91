"""The FP8 e4m3 storage form of the logistic model: X stored as OCP e4m3fn
bytes with a power-of-two scale per column (X_q[n][d] = e4m3(code) *
scale[d]), read by the VALU tile kernel with fp32 arithmetic.

* The library's quantiser (nax_fp8_quantize) is pinned against torch's own
  float8_e4m3fn conversion (round to nearest even), an implementation that
  shares nothing with it, and the scales against their defining rule (the
  power of two s >= max|X[:, d]| / 448).
* U and dU/dz of the fp8 model against the float64 oracle evaluated on the
  dequantised design X_q (arithmetic error only), at sizes of the
  configurations the form is meant for, and the error against the original
  fp32 X reported (quantisation of X: a different model, not an error)."""
import json
import sys

import numpy as np
import pytest
import torch

from oracle import models as om

pytestmark = pytest.mark.gpu


def _quantize(X):
    from numpyro_amd import _lib
    N, D = X.shape
    q = torch.empty((N, D), dtype=torch.uint8, device="cuda")
    s = torch.empty(D, dtype=torch.float32, device="cuda")
    _lib.check(_lib.load().nax_fp8_quantize(_lib.ptr(X), N, D, _lib.ptr(q), _lib.ptr(s)),
               "nax_fp8_quantize")
    torch.cuda.synchronize()
    return q, s


@pytest.mark.parametrize("N,D", [(5000, 7), (20000, 55), (3000, 256)])
def test_fp8_quantiser_matches_torch_float8_e4m3fn(N, D):
    rng = np.random.default_rng(D)
    X = rng.normal(size=(N, D)).astype(np.float32)
    X[:, : D // 5] *= 10.0
    X[:, -1] = 0.0                                   # an all-zero column: scale 1
    X[0, 1] = 1e-6                                   # below the subnormal range
    Xt = torch.from_numpy(X).cuda()
    q, s = _quantize(Xt)
    s = s.cpu().numpy()
    m = np.abs(X).max(0)
    for d in range(D):
        if m[d] == 0:
            assert s[d] == 1.0
        else:
            assert s[d] >= m[d] / 448 and s[d] / 2 < m[d] / 448
            assert np.log2(s[d]) == np.round(np.log2(s[d]))
    want = (torch.from_numpy(X) / torch.from_numpy(s)).to(torch.float8_e4m3fn)
    np.testing.assert_array_equal(q.cpu().numpy(), want.view(torch.uint8).numpy())


@pytest.mark.parametrize("N,D,C", [(4000, 26, 4), (50000, 55, 64), (20000, 256, 32)])
def test_fp8_potential_grad_vs_oracle(N, D, C):
    from numpyro_amd import _lib
    from numpyro_amd.infer.hmc import Engine
    from numpyro_amd.models import LogisticRegression
    rng = np.random.default_rng(N + D)
    X = rng.normal(size=(N, D)).astype(np.float32)
    X[:, : D // 5] *= 10.0
    beta = 0.1 * rng.normal(size=D)
    y = (rng.uniform(size=N) < 1 / (1 + np.exp(-X.astype(np.float64) @ beta))).astype(np.float32)
    Xt = torch.from_numpy(X).cuda()
    q, s = _quantize(Xt)
    Xq = (q.cpu().view(torch.float8_e4m3fn).to(torch.float32) * s.cpu()).numpy()
    eng = Engine(LogisticRegression(X, y, x_dtype="fp8"), C, "NUTS", 0)
    ld = _lib.stride(eng.h)
    Z = (beta + 0.01 * rng.normal(size=(C, D))).astype(np.float32)
    zp = torch.zeros((C, ld), device="cuda")
    zp[:, :D] = torch.from_numpy(Z).cuda()
    pe = torch.empty(C, dtype=torch.float64, device="cuda")
    g = torch.zeros((C, ld), device="cuda")
    _lib.check(_lib.load().nax_potential_grad(eng.h, _lib.ptr(zp), ld, C, _lib.ptr(pe),
                                              _lib.ptr(g)), "potential_grad")
    pe = pe.cpu().numpy()
    g = g[:, :D].cpu().numpy()
    du = dg = du_x = dg_x = 0.0
    for c in range(C):
        z64 = Z[c].astype(np.float64)
        u, gr = om.logistic_potential_grad(z64, Xq.astype(np.float64), y.astype(np.float64))
        u0, g0 = om.logistic_potential_grad(z64, X.astype(np.float64), y.astype(np.float64))
        du = max(du, abs(pe[c] - u) / abs(u))
        dg = max(dg, np.linalg.norm(g[c] - gr) / np.linalg.norm(gr))
        du_x = max(du_x, abs(pe[c] - u0) / abs(u0))
        dg_x = max(dg_x, np.linalg.norm(g[c] - g0) / np.linalg.norm(g0))
    rec = {"form": "fp8 e4m3 + pow2 column scales", "N": N, "D": D, "chains": C,
           "U_rel_err_vs_quantised_X": du, "grad_rel_err_vs_quantised_X": dg,
           "U_rel_err_vs_fp32_X": du_x, "grad_rel_err_vs_fp32_X": dg_x}
    print(json.dumps(rec), file=sys.stderr, flush=True)
    assert du < 1e-5 and dg < 1e-4, rec
