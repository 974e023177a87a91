"""The hand-written adjoint kernels behind torch.autograd.Function
(`numpyro_amd.models.potential_energy`): forward values and backward
gradients of every built-in model against the float64 oracle, whose
gradients are pinned to central finite differences in
tests/test_oracle_models.py.  U: 1e-5 relative (fp32 data, fp64 sums);
gradients: 1e-4 relative to the gradient's scale."""
import numpy as np
import pytest
import torch

from oracle import models as om

pytestmark = pytest.mark.gpu


def _check(model, Z, oracle_vg, g_rtol=1e-4):
    from numpyro_amd.models import potential_energy
    z = torch.tensor(Z, device="cuda", requires_grad=True)
    u = potential_energy(model, z)
    w = torch.linspace(0.5, 1.5, Z.shape[0], dtype=torch.float64, device="cuda")
    (u * w).sum().backward()
    g = z.grad.cpu().numpy().astype(np.float64)
    uc = u.detach().cpu().numpy()
    for c in range(Z.shape[0]):
        ou, og = oracle_vg(c, Z[c].astype(np.float64))
        np.testing.assert_allclose(uc[c], ou, rtol=1e-5)
        scale = np.abs(og).max()
        np.testing.assert_allclose(g[c, :og.size] / float(w[c]), og, rtol=0, atol=g_rtol * scale)


def test_autograd_logistic():
    from numpyro_amd.models import LogisticRegression
    rng = np.random.default_rng(0)
    N, D, C = 3000, 17, 6
    X = rng.normal(size=(N, D)).astype(np.float32)
    y = (rng.uniform(size=N) < 0.4).astype(np.float32)
    Z = rng.normal(scale=0.3, size=(C, D)).astype(np.float32)
    _check(LogisticRegression(X, y), Z, lambda c, z: om.logistic_potential_grad(z, X, y, 1.0))


@pytest.mark.parametrize("noncentred", [False, True])
def test_autograd_sv(noncentred):
    from numpyro_amd.models import StochasticVolatility
    from numpyro_amd.workloads import make_sv
    r, _ = make_sv(T=200, seed=3)
    rng = np.random.default_rng(1)
    C, T = 4, 200
    Z = rng.normal(scale=0.2, size=(C, T + 2)).astype(np.float32)
    Z[:, 0] = np.log(5.0)
    f = om.sv_nc_potential_grad if noncentred else om.sv_potential_grad
    _check(StochasticVolatility(r, noncentred=noncentred), Z,
           lambda c, z: f(z, r.astype(np.float64)))


def test_autograd_hier():
    from numpyro_amd.models import HierarchicalLogistic
    rng = np.random.default_rng(2)
    shapes = [(300, 5), (77, 12), (1000, 3)]
    Xs = [rng.normal(size=s).astype(np.float32) for s in shapes]
    ys = [(rng.uniform(size=s[0]) < 0.5).astype(np.float32) for s in shapes]
    K = 2
    m = HierarchicalLogistic(Xs, ys, chains_per_group=K)
    C = len(shapes) * K
    Z = np.zeros((C, m.dim), np.float32)
    for c in range(C):
        d = shapes[c // K][1]
        Z[c, :d + 2] = rng.normal(scale=0.3, size=d + 2)

    def vg(c, z):
        i = c // K
        d = shapes[i][1]
        return om.hier_potential_grad(z[:d + 2], Xs[i], ys[i])
    _check(m, Z, vg)
