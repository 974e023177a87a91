"""1024-chain NUTS on a design shaped like the 2M x 256 benchmark (D=256,
X ~ N(0,1) with 51 columns scaled x10, beta* ~ N(0, 0.1^2); N reduced to
200,000 rows so that the runs finish within the test budget).  Each run
takes its own warmup from the same Laplace initialisation (init_params +
inverse_mass_matrix, find_heuristic_step_size=True, as the benchmark does).

The BF16-MFMA form reads a bf16 (round-to-nearest-even) copy of X, so it
samples the posterior of the bf16-rounded design.  It is checked against the
plain FP32 form run on that same rounded design (fp32 copy of the bf16
values: the same target, only the arithmetic differs):

* both mix: max split R-hat over the 256 coordinates < 1.01 (1024 chains);
* both rarely diverge: fewer than 1 % of the sampling transitions;
* they agree: posterior means within 3 MCSE of each other (MCSE from each
  run's own ESS) for at least 98 % of coordinates, within 4 MCSE for all
  (256 coordinates: a 3-MCSE bound on every one would fail by chance about
  half the time for two independent runs of one target).

The shift of the target itself (bf16-rounded X against the original fp32
X, a third run) is reported in posterior standard deviations and bounded by
0.05 sd: with ESS ~1e5 it is many MCSE, so it is not an MCSE test."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _run(X, y, x_dtype, z0, imm, keys, warmup, samples):
    from numpyro_amd.infer.hmc import Engine
    from numpyro_amd.models import LogisticRegression
    model = LogisticRegression(X, y, x_dtype=x_dtype, kernel="mfma" if x_dtype == "fp32" else
                               "auto")
    C = z0.shape[0]
    eng = Engine(model, C, "NUTS", warmup, find_heuristic_step_size=True)
    eng.init(z0, keys, imm)
    eng.run(warmup)
    out = eng.run(warmup + samples, collect_fields=("z", "diverging", "num_steps"),
                  start_idx=warmup, size=samples)
    torch.cuda.synchronize()
    return out


def test_c5_shaped_1024_chains_bf16_vs_fp32():
    from numpyro_amd import random as nrandom
    from numpyro_amd import workloads
    from numpyro_amd.diagnostics import device_diagnostics
    N, C, W, S = 200_000, 1024, 150, 150
    print(f"[c5-shaped] N={N} D=256 chains={C} warmup={W} samples={S}", flush=True)
    X, y, beta, cfg = workloads.make("c5", N=N, chains=C)
    Xq = X.to(torch.bfloat16).to(torch.float32).contiguous()  # the bf16 form's design
    keys = nrandom.chain_init_keys(nrandom.PRNGKey(0), C)
    z0, imm = workloads.laplace_init(X, y, C)
    res = {}
    for name, xd, design in (("bf16", "bf16", X), ("fp32 on bf16-rounded X", "fp32", Xq),
                             ("fp32 on fp32 X", "fp32", X)):
        out = _run(design, y, xd, z0, imm, keys, W, S)
        z = out["z"]
        ess, rhat = device_diagnostics(z)
        ess = ess.cpu().numpy()
        rhat = rhat.cpu().numpy()
        zf = z.double()
        mean = zf.mean((0, 1)).cpu().numpy()
        sd = zf.reshape(-1, z.shape[-1]).std(0).cpu().numpy()
        div = int(out["diverging"].sum())
        print(f"[c5-shaped] {name}: max R-hat {rhat.max():.4f}, min ESS {ess.min():.0f}, "
              f"divergences {div}, mean tree {float(out['num_steps'].float().mean()):.1f}",
              flush=True)
        assert rhat.max() < 1.01, (name, rhat.max())
        assert div < 0.01 * C * S, (name, div)
        res[name] = (mean, sd, ess)
    m16, s16, e16 = res["bf16"]
    mq, sq, eq = res["fp32 on bf16-rounded X"]
    m32, s32, _ = res["fp32 on fp32 X"]
    mcse = np.sqrt(s16 ** 2 / e16 + sq ** 2 / eq)
    zs = np.abs(m16 - mq) / mcse
    shift = np.abs(mq - m32) / s32
    print(f"[c5-shaped] |bf16 - fp32| / MCSE (same rounded X): max {zs.max():.2f}, "
          f"within 3: {np.mean(zs < 3):.3f}; target shift of bf16 rounding: max "
          f"{shift.max():.4f} sd, mean {shift.mean():.4f} sd", flush=True)
    assert np.mean(zs < 3) >= 0.98, zs
    assert zs.max() < 4, zs
    assert shift.max() < 0.05, shift
