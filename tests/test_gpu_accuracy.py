"""Error of U and dU/dz against float64 for every benchmark configuration.

    python -m pytest tests/test_gpu_accuracy.py -m gpu -s   # prints one JSON line each

For each configuration the engine's potential + adjoint kernels are
evaluated at fixed positions near the data-generating values and compared
with the same potential computed in float64 (torch float64 on the GPU for the
logistic designs, numpy float64 for the others): max over chains of
|U - U64| / |U64| and of ||g - g64|| / ||g64||.  The bf16 form is compared
both with the float64 potential of the original fp32 X (quantisation of X
included) and of its bf16-rounded copy (arithmetic error only).
"""
import json
import sys

import numpy as np
import pytest
import torch

from oracle import models as om

pytestmark = pytest.mark.gpu


def _rel(u, u64, g, g64):
    du = float(np.max(np.abs(u - u64) / np.abs(u64)))
    dg = float(np.max(np.linalg.norm(g - g64, axis=1) / np.linalg.norm(g64, axis=1)))
    return du, dg


def _logistic64(X, y, Z):
    """float64 U and grad on the GPU for positions Z [C, D] (prior N(0, 1))."""
    X64 = X.to(torch.float64)
    y64 = y.to(torch.float64)
    Z64 = Z.to(torch.float64)
    L = X64 @ Z64.T                                    # [N, C]
    u = (torch.nn.functional.softplus(L) - y64[:, None] * L).sum(0)
    u = u + 0.5 * (Z64 * Z64).sum(1) + Z.shape[1] * 0.5 * np.log(2 * np.pi)
    R = torch.sigmoid(L) - y64[:, None]
    g = (X64.T @ R).T + Z64
    return u.cpu().numpy(), g.cpu().numpy()


def logistic(name, forms, chains=None):
    from numpyro_amd import workloads
    from numpyro_amd.models import LogisticRegression, _device_value_and_grad
    X, y, beta, cfg = workloads.make(name)
    C = chains or cfg["chains"]
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    Z = (beta[None, :] + 0.01 * torch.randn((C, cfg["D"]), generator=gen, device="cuda"))
    Z = Z.float().contiguous()
    u64, g64 = _logistic64(X, y, Z)
    out = []
    for form in forms:
        # "bf16": one copy of X, bf16 residuals in X^T R (the default);
        # "bf16-two-copy": an fp16 copy of X for fp16 residuals
        m = (LogisticRegression(X, y, x_dtype="bf16", kernel="two-copy")
             if form == "bf16-two-copy" else LogisticRegression(X, y, x_dtype=form))
        u, g = _device_value_and_grad(m, Z)
        u, g = u.cpu().numpy(), g.cpu().numpy().astype(np.float64)
        du, dg = _rel(u, u64, g, g64)
        rec = {"config": name, "form": form, "chains": C, "N": cfg["N"], "D": cfg["D"],
               "U_rel_err": du, "grad_rel_err": dg}
        if form.startswith("bf16"):
            ub, gb = _logistic64(X.to(torch.bfloat16).float(), y, Z)
            rec["U_rel_err_vs_rounded_X"], rec["grad_rel_err_vs_rounded_X"] = _rel(u, ub, g, gb)
        out.append(rec)
        del m
    return out


def sv(noncentred):
    from numpyro_amd.models import StochasticVolatility, _device_value_and_grad
    from numpyro_amd.workloads import make_sv
    r, meta = make_sv(3000, 3)
    C, T = 128, 3000
    rng = np.random.default_rng(11)
    Z = np.zeros((C, T + 2), np.float32)
    if noncentred:
        Z[:, 0] = np.log(max(meta["nu"], 1.0))
        Z[:, 1] = np.log(meta["sigma"])
        zeta = np.diff(np.concatenate([[0.0], meta["s"]])) / meta["sigma"]
        Z[:, 2:] = zeta + 0.05 * rng.normal(size=(C, T))
    else:
        Z[:, 0] = np.log(max(meta["nu"], 1.0))
        Z[:, 1:T + 1] = meta["s"] + 0.05 * rng.normal(size=(C, T))
        Z[:, T + 1] = np.log(meta["sigma"])
    m = StochasticVolatility(r, noncentred=noncentred)
    u, g = _device_value_and_grad(m, torch.tensor(Z))
    u, g = u.cpu().numpy(), g.cpu().numpy().astype(np.float64)
    f = om.sv_nc_potential_grad if noncentred else om.sv_potential_grad
    ref = [f(Z[c].astype(np.float64), r.astype(np.float64)) for c in range(C)]
    du, dg = _rel(u, np.array([a for a, _ in ref]), g, np.stack([b for _, b in ref]))
    return [{"config": "c3", "form": "fp32", "noncentred": noncentred, "chains": C, "T": T,
             "U_rel_err": du, "grad_rel_err": dg}]


def hier():
    from numpyro_amd.models import HierarchicalLogistic, _device_value_and_grad
    from numpyro_amd.workloads import make_hier
    Xs, ys = make_hier(1024, 4)
    K = 4
    m = HierarchicalLogistic(Xs, ys, chains_per_group=K)
    C = len(Xs) * K
    rng = np.random.default_rng(12)
    Z = np.zeros((C, m.dim), np.float32)
    for c in range(C):
        d = Xs[c // K].shape[1]
        Z[c, :d + 2] = 0.3 * rng.normal(size=d + 2)
    u, g = _device_value_and_grad(m, torch.tensor(Z))
    u, g = u.cpu().numpy(), g.cpu().numpy().astype(np.float64)
    du, dg = 0.0, 0.0
    for c in range(C):
        i = c // K
        d = Xs[i].shape[1]
        ou, og = om.hier_potential_grad(Z[c, :d + 2].astype(np.float64), Xs[i], ys[i])
        du = max(du, abs(u[c] - ou) / abs(ou))
        dg = max(dg, float(np.linalg.norm(g[c, :d + 2] - og) / np.linalg.norm(og)))
    return [{"config": "c4", "form": "fp32", "chains": C, "U_rel_err": du, "grad_rel_err": dg}]


def _report(recs):
    for r in recs:
        print(json.dumps(r), flush=True)
        print(f"[accuracy] {r}", file=sys.stderr, flush=True)


# Bounds: fp32 data with fp32 products and fp64 energy sums; the bf16 forms'
# z enters as a bf16 hi+lo pair (~16 bits), their residuals as bf16 (one
# copy of X) or fp16 (two copies).
def test_accuracy_logistic_c1_c2():
    recs = logistic("c1", ["fp32"]) + logistic("c2", ["fp32", "bf16", "bf16-two-copy"])
    _report(recs)
    for r in recs:
        # against the fp32 X the bf16 form also carries the rounding of X
        # itself (a different model, not an arithmetic error): loose bounds
        assert r["U_rel_err"] < (1e-5 if r["form"] == "fp32" else 1e-2), r
        assert r["grad_rel_err"] < (1e-4 if r["form"] == "fp32" else 0.1), r
        if r["form"].startswith("bf16"):
            assert r["U_rel_err_vs_rounded_X"] < 1e-5, r
            assert r["grad_rel_err_vs_rounded_X"] < 5e-3, r


@pytest.mark.parametrize("noncentred", [False, True])
def test_accuracy_sv_c3(noncentred):
    recs = sv(noncentred)
    _report(recs)
    assert recs[0]["U_rel_err"] < 1e-5 and recs[0]["grad_rel_err"] < 1e-4, recs


def test_accuracy_hier_c4():
    recs = hier()
    _report(recs)
    assert recs[0]["U_rel_err"] < 1e-5 and recs[0]["grad_rel_err"] < 1e-4, recs


def test_accuracy_logistic_c5():
    recs = logistic("c5", ["fp32", "bf16", "bf16-two-copy"], chains=64)
    _report(recs)
    for r in recs:
        # against the fp32 X the bf16 forms also carry the rounding of X
        # itself (a different model, not an arithmetic error): loose bounds
        assert r["U_rel_err"] < (1e-5 if r["form"] == "fp32" else 1e-2), r
        assert r["grad_rel_err"] < (1e-4 if r["form"] == "fp32" else 0.1), r
        if r["form"].startswith("bf16"):
            assert r["U_rel_err_vs_rounded_X"] < 1e-5, r
            assert r["grad_rel_err_vs_rounded_X"] < 5e-3, r


def test_energy_differences_c5():
    # Accept/reject and the tree weights use differences of energies between
    # nearby positions (hmc_util.py:866-874), not U itself: at the full C5
    # design (N = 2M, |U| ~ 1e6) the error of U(z') - U(z) against float64
    # (on the same design, bf16-rounded for the bf16 form) for steps of the
    # size of a leapfrog move must stay far below one nat.  Positions are
    # drawn where the chains are: the Laplace mode plus one Laplace standard
    # deviation of Gaussian noise per coordinate; steps are half a standard
    # deviation.  (Far outside the posterior the gradient is ~100x larger
    # and so is the energy error of the bf16 form's ~2^-17 position
    # rounding: U errors of ~0.4 nats at 0.01 from beta*, tens of sds.)
    from numpyro_amd import workloads
    from numpyro_amd.models import LogisticRegression, _device_value_and_grad
    X, y, beta, cfg = workloads.make("c5")
    C = 32
    gen = torch.Generator(device="cuda")
    gen.manual_seed(9)
    mode, var = workloads.laplace(X, y)
    sd = var.sqrt()
    noise = torch.randn((C, cfg["D"]), generator=gen, device="cuda", dtype=torch.float64)
    Z = (mode[None, :] + sd[None, :] * noise).float().contiguous()
    step = torch.randn((C, cfg["D"]), generator=gen, device="cuda", dtype=torch.float64)
    Z2 = (Z.double() + 0.5 * sd[None, :] * step).float().contiguous()
    recs = []
    for form in ("fp32", "bf16"):
        m = LogisticRegression(X, y, x_dtype=form, kernel="mfma" if form == "fp32" else "auto")
        u1, _ = _device_value_and_grad(m, Z)
        u2, _ = _device_value_and_grad(m, Z2)
        Xr = X.to(torch.bfloat16).float() if form == "bf16" else X
        a, _ = _logistic64(Xr, y, Z)
        b, _ = _logistic64(Xr, y, Z2)
        d = (u2 - u1).cpu().numpy()
        d64 = b - a
        err = np.abs(d - d64)
        rec = {"config": "c5", "form": form, "chains": C, "max_abs_U": float(np.abs(a).max()),
               "mean_abs_dU": float(np.abs(d64).mean()), "max_abs_dU_err": float(err.max()),
               "max_abs_U_err": float(np.abs(u1.cpu().numpy() - a).max())}
        recs.append(rec)
        del m
    _report(recs)
    for r in recs:
        assert r["max_abs_dU_err"] < 0.05, r
