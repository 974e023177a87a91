"""On-device ESS / split R-hat (nax_diagnostics) against the host NumPy
formulas of numpyro_amd.diagnostics, which are pinned to the reference's
known-answer tests (test/test_diagnostics.py:61-108) in
tests/test_oracle_random_diag.py.  Direct lag sums vs FFT: agreement to
float64 rounding."""
import numpy as np
import pytest
import torch

from numpyro_amd import diagnostics as nd

pytestmark = pytest.mark.gpu


def _ar1(C, S, P, phi, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((C, S, P))
    e = rng.normal(size=(C, S, P))
    x[:, 0] = e[:, 0]
    for t in range(1, S):
        x[:, t] = phi * x[:, t - 1] + e[:, t]
    x += rng.normal(size=(C, 1, P)) * 0.1  # a little between-chain spread
    return x.astype(np.float32)


@pytest.mark.parametrize("C, S, P, phi", [(4, 1000, 3, 0.9), (1, 2, 1, 0.0), (2, 3, 5, 0.5),
                                          (16, 120, 256, 0.6), (3, 4096, 2, 0.99),
                                          (1024, 120, 4, 0.3), (7, 501, 9, -0.4)])
def test_device_diagnostics_match_host(C, S, P, phi):
    x = _ar1(C, S, P, phi, C * 1000 + S + P)
    ess, rhat = nd.device_diagnostics(torch.from_numpy(x))
    want_ess = nd.effective_sample_size(x.astype(np.float64))
    np.testing.assert_allclose(ess.cpu().numpy(), want_ess, rtol=1e-8)
    if S // 2 >= 2:
        want_rhat = nd.split_gelman_rubin(x.astype(np.float64))
        np.testing.assert_allclose(rhat.cpu().numpy(), want_rhat, rtol=1e-9)
    else:
        assert np.isnan(rhat.cpu().numpy()).all()


def test_device_diagnostics_refuses_long_chains():
    with pytest.raises(ValueError):
        nd.device_diagnostics(torch.zeros((1, 4097, 1)))
