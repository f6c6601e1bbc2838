"""Seeded synthetic wave functions for the parity tests (test helper)."""
import math

import numpy as np


def grf(n: int, seed: int, slope: float = -2.5, rms: float = 1.0) -> np.ndarray:
    """Zero-mean Gaussian random field on n^3 with P(k) ~ k^slope up to the
    Nyquist frequency, scaled to the given rms."""
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((n, n, n))
    k = np.fft.fftfreq(n) * n
    kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2)
    amp = np.zeros_like(kk)
    nz = kk > 0
    amp[nz] = kk[nz] ** (slope / 2)
    amp[kk > n / 2] = 0.0
    f = np.fft.ifftn(np.fft.fftn(white) * amp).real
    f -= f.mean()
    s = f.std()
    return f * (rms / s) if s > 0 else f


def make_psi(n: int, seed: int, mfdm_solar: float, rhoc: float = 3.51059e10, delta_rms: float = 0.05,
             phase_rms: float = 1.0) -> np.ndarray:
    """psi = sqrt(rho / m_s) exp(i theta) with rho = rhoc (1 + delta)."""
    delta = grf(n, seed, rms=delta_rms)
    theta = grf(n, seed + 7919, rms=phase_rms)
    rho = rhoc * (1.0 + np.clip(delta, -0.9, None))
    return np.sqrt(rho / mfdm_solar) * np.exp(1j * theta)


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
