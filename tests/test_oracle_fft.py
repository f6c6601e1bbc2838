"""Pins the oracle's FFTs: direct DFT summation, numpy's pocketfft, Parseval
and the impulse response.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O


def direct_dft3(a: np.ndarray, sign: int) -> np.ndarray:
    """X[k] = sum_n a[n] exp(sign 2 pi i k.n / N), evaluated term by term
    along each axis from the defining formula."""
    n = a.shape[0]
    idx = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(idx, idx) / n)
    return np.einsum("ia,jb,kc,abc->ijk", W, W, W, a)


@pytest.mark.parametrize("n", [4, 8])
@pytest.mark.parametrize("sign", [-1, 1])
def test_c2c_direct_dft(n, sign):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
    ref = direct_dft3(a, sign)
    out = O.fft3d_c2c(a, sign)
    assert np.abs(out - ref).max() / np.abs(ref).max() < 1e-14


@pytest.mark.parametrize("n", [16, 32])
def test_c2c_numpy(n):
    rng = np.random.default_rng(3)
    a = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
    assert np.abs(O.fft3d_c2c(a, -1) - np.fft.fftn(a)).max() / np.abs(np.fft.fftn(a)).max() < 1e-14
    assert np.abs(O.fft3d_c2c(a, +1) - np.fft.ifftn(a) * n ** 3).max() / (np.abs(a).max() * n ** 1.5) < 1e-13


@pytest.mark.parametrize("n", [8, 16, 32])
def test_r2c_c2r(n):
    rng = np.random.default_rng(5)
    r = rng.standard_normal((n, n, n))
    h = O.fft3d_r2c(r)
    assert h.shape == (n, n, n // 2 + 1)
    assert np.abs(h - np.fft.rfftn(r)).max() / np.abs(h).max() < 1e-14
    back = O.fft3d_c2r(h)
    assert np.abs(back / n ** 3 - r).max() < 1e-13
    # c2r of an arbitrary half spectrum equals numpy's irfftn (unnormalised)
    hh = rng.standard_normal(h.shape) + 1j * rng.standard_normal(h.shape)
    assert np.abs(O.fft3d_c2r(hh) - np.fft.irfftn(hh, s=(n, n, n), axes=(0, 1, 2)) * n ** 3).max() < 1e-10 * n ** 3


def test_parseval_and_impulse():
    n = 16
    rng = np.random.default_rng(9)
    a = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
    f = O.fft3d_c2c(a, -1)
    assert np.sum(np.abs(f) ** 2) / n ** 3 == pytest.approx(np.sum(np.abs(a) ** 2), rel=1e-13)
    d = np.zeros((n, n, n), complex)
    d[0, 0, 0] = 1.0
    assert np.abs(O.fft3d_c2c(d, -1) - 1.0).max() < 1e-15
    # a shifted impulse gives a pure plane wave
    d[:] = 0
    d[1, 2, 3] = 1.0
    kx, ky, kz = np.meshgrid(*[np.arange(n)] * 3, indexing="ij")
    pw = np.exp(-2j * np.pi * (kx * 1 + ky * 2 + kz * 3) / n)
    assert np.abs(O.fft3d_c2c(d, -1) - pw).max() < 1e-14
