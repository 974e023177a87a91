"""Pins oracle.jax_random (published threefry KATs, jax's documented split
value) and oracle.diagnostics (test/test_diagnostics.py KATs)."""
import numpy as np
from numpy.testing import assert_allclose
import pytest
from scipy.fftpack import next_fast_len

from oracle import diagnostics as dg
from oracle import jax_random as jr


@pytest.mark.parametrize("key, ctr, expected", [
    ((0, 0), (0, 0), (0x6B200159, 0x99BA4EFE)),
    ((0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF), (0x1CB996FC, 0xBB002BE7)),
    ((0x13198A2E, 0x03707344), (0x243F6A88, 0x85A308D3), (0xC4923A9C, 0x483DF7A0)),
])
def test_threefry_kat(key, ctr, expected):
    # Random123 kat_vectors for threefry2x32_20 (also jax's tests/random_test.py)
    o0, o1 = jr.threefry2x32(key, [ctr[0]], [ctr[1]])
    assert (int(o0[0]), int(o1[0])) == expected


def test_split_original_known_value():
    # jax documentation: random.split(PRNGKey(0)) in the original derivation
    k = jr.split(jr.PRNGKey(0), 2, partitionable=False)
    assert k.tolist() == [[4146024105, 967050713], [2718843009, 1272950319]]


def test_uniform_mantissa_trick():
    bits = np.array([0, 0xFFFFFFFF, 0x80000000], dtype=np.uint32)
    f = jr.bits_to_unit_float32(bits)
    assert f[0] == 0.0 and f[1] == np.float32(1 - 2 ** -23) and f[2] == 0.5


def test_normal_moments():
    x = jr.normal(jr.PRNGKey(3), 200000)
    assert abs(x.mean()) < 0.01 and abs(x.std() - 1) < 0.01


def test_partitionable_split_is_threefry_of_iota():
    key = np.array([7, 9], np.uint32)
    ks = jr.split(key, 3)
    for i in range(3):
        o0, o1 = jr.threefry2x32(key, [0], [i])
        assert ks[i].tolist() == [int(o0[0]), int(o1[0])]


@pytest.mark.parametrize("target", [433, 124, 25, 300, 1, 3, 7])
def test_fft_next_fast_len(target):
    assert dg.fft_next_fast_len(target) == next_fast_len(target)


def test_autocorrelation():
    x = np.arange(10.0)
    expected = np.array([1, 0.78, 0.52, 0.21, -0.13, -0.52, -0.94, -1.4, -1.91, -2.45])
    assert_allclose(dg.autocorrelation(x, bias=False), expected, atol=0.01)
    assert_allclose(dg.autocorrelation(x, bias=True),
                    expected * np.arange(10, 0.0, -1) / 10, atol=0.01)


def test_autocovariance():
    expected = np.array([8.25, 6.42, 4.25, 1.75, -1.08, -4.25, -7.75, -11.58, -15.75, -20.25])
    assert_allclose(dg.autocovariance(np.arange(10.0), bias=False), expected, atol=0.01)


def test_gelman_rubin():
    x = np.stack([np.arange(10.0), np.arange(10.0) + 1])
    assert_allclose(dg.gelman_rubin(x), 0.98, atol=0.01)


def test_split_gelman_rubin():
    x = np.random.RandomState(0).normal(size=(2, 10))
    assert_allclose(dg.gelman_rubin(x.reshape(4, 5)), dg.split_gelman_rubin(x))


def test_effective_sample_size():
    x = np.arange(1000.0).reshape(100, 10)
    assert_allclose(dg.effective_sample_size(x, bias=False), 52.64, atol=0.01)


def test_ess_iid_close_to_n():
    x = np.random.RandomState(1).normal(size=(4, 2000))
    ess = dg.effective_sample_size(x)
    assert 0.8 * 8000 < ess < 1.2 * 8000
