"""Seeded synthetic wave functions for the parity tests (test helper)."""
import math
import os

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


def snapshot_bytes(psi: np.ndarray, xmax: float, a: float, mfdm_ev: float, g_ev_cm3: float) -> bytes:
    """A snapshot in the reference's layout (src/write3D.h:86-97,
    src/readsnap.cpp:48-52): f64 mfdm [eV], f64 g [eV cm^3], f64 a,
    f64 x[N] with x_i = xmax/N - xmax + i dx (src/struct3D.h:72), then psi
    as N^3 complex128 in C order."""
    n = psi.shape[0]
    x = np.array([xmax / n - xmax + i * (2.0 * xmax / n) for i in range(n)], "<f8")
    head = np.array([mfdm_ev, g_ev_cm3, a], "<f8")
    return head.tobytes() + x.tobytes() + np.ascontiguousarray(psi, "<c16").tobytes()


# inputs of the golden snapshot fixture (tests/golden/readsnap_n8_z3.txt)
GOLDEN_SNAP = {"n": 8, "seed": 3, "xmax_mpc_h": 5.0, "a": 0.1, "mfdm_ev": 2.5e-24, "g_ev_cm3": 1e-60, "zslice": 3}


def _upsample_axis(a: np.ndarray, n: int, axis: int) -> np.ndarray:
    """Band-limited interpolation along one axis: the m-point spectrum
    zero-padded to n points (the Nyquist term split between +m/2 and -m/2,
    so the result passes through the m input points)."""
    import scipy.fft as sf
    m = a.shape[axis]
    w = min(16, os.cpu_count() or 1)
    A = sf.fft(a, axis=axis, workers=w)
    shape = list(a.shape)
    shape[axis] = n
    B = np.zeros(shape, np.complex128)
    lo = [slice(None)] * a.ndim
    hi = [slice(None)] * a.ndim
    lo[axis] = slice(0, m // 2)
    hi[axis] = slice(n - (m // 2 - 1), n)
    src_hi = [slice(None)] * a.ndim
    src_hi[axis] = slice(m - (m // 2 - 1), m)
    B[tuple(lo)] = A[tuple(lo)]
    B[tuple(hi)] = A[tuple(src_hi)]
    nyq_src = [slice(None)] * a.ndim
    nyq_src[axis] = slice(m // 2, m // 2 + 1)
    for dst in (m // 2, n - m // 2):
        d = [slice(None)] * a.ndim
        d[axis] = slice(dst, dst + 1)
        B[tuple(d)] = 0.5 * A[tuple(nyq_src)]
    del A
    out = sf.ifft(B, axis=axis, workers=w, overwrite_x=True)
    out *= n / m
    return out


def smooth_psi(n: int, m: int, seed: int, mfdm_solar: float, **kw) -> np.ndarray:
    """make_psi at m^3, interpolated spectrally to n^3: a smooth field at
    the large mesh sizes without an n^3 Gaussian random field."""
    a = make_psi(m, seed, mfdm_solar, **kw)
    for ax in range(3):
        a = _upsample_axis(a, n, ax)
    return a
