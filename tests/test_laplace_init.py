"""The Laplace initialisation the benchmark and the 1024-chain test use
(numpyro_amd.workloads.laplace): the mode is a stationary point of U and the
variances are the diagonal of the inverse of U's Hessian, both checked
against numpy computations that do not share the library's code (an
analytic gradient and a central finite-difference Hessian)."""
import numpy as np
import torch


def _grad_u(b, X, y):
    p = 1.0 / (1.0 + np.exp(-X @ b))
    return X.T @ (p - y) + b


def test_laplace_mode_and_variances():
    from numpyro_amd.workloads import laplace, laplace_init
    rng = np.random.default_rng(0)
    N, D = 4000, 6
    X = rng.normal(size=(N, D))
    X[:, 2] *= 10.0
    beta = 0.3 * rng.normal(size=D)
    y = (rng.uniform(size=N) < 1 / (1 + np.exp(-X @ beta))).astype(np.float64)
    mode, var = laplace(torch.from_numpy(X), torch.from_numpy(y), chunk=1000)
    mode = mode.numpy()
    assert np.abs(_grad_u(mode, X, y)).max() < 1e-8
    h = 1e-5
    H = np.stack([(_grad_u(mode + h * e, X, y) - _grad_u(mode - h * e, X, y)) / (2 * h)
                  for e in np.eye(D)])
    np.testing.assert_allclose(var.numpy(), np.diag(np.linalg.inv(H)), rtol=1e-5)
    z0, imm = laplace_init(torch.from_numpy(X).float(), torch.from_numpy(y).float(), 5)
    assert z0.shape == (5, D) and imm.shape == (D,)
    sd = np.sqrt(imm.numpy())
    assert np.all(np.abs(z0.numpy() - mode) <= 2 * sd * (1 + 1e-4) + 1e-6)
