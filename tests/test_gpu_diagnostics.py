"""GPU: the matter power spectrum diagnostic (no counterpart in the
reference) against a single-cosine closed form and numpy's rfftn."""
import numpy as np
import pytest

from tests.fields import make_psi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [0, 1])
def test_single_mode_power(prec):
    import ifdm_code_amd as F
    n = 32
    cfg = F.Config.from_cli(2.5e-24, 0.0, 5.0, n, precision=prec)
    i, j, k = np.meshgrid(*[np.arange(n)] * 3, indexing="ij")
    m = (3, 0, 4)                                 # |n| = 5
    eps = 0.02
    rho = cfg.rhoc * (1 + eps * np.cos(2 * np.pi * (m[0] * i + m[1] * j + m[2] * k) / n))
    psi = np.sqrt(rho / cfg.mfdm_solar) + 0j
    with F.Solver(cfg) as s:
        s.set_psi(psi)
        kk, pk, nm = s.power_spectrum(n // 2)
    L = 2 * cfg.xmax
    tot = pk * nm
    assert tot[4] == pytest.approx(L ** 3 * eps ** 2 / 2, rel=1e-10 if prec == 0 else 1e-5)
    others = np.delete(tot, 4)
    assert np.abs(others).max() < (1e-20 if prec == 0 else 1e-9) * tot[4]
    # the shell |n| in [4.5, 5.5): its mode-averaged |k|
    f = np.fft.fftfreq(n) * n
    nz = np.arange(n // 2 + 1)
    nn = np.sqrt(f[:, None, None] ** 2 + f[None, :, None] ** 2 + nz[None, None, :] ** 2)
    w = np.where((nz == 0) | (nz == n // 2), 1.0, 2.0)[None, None, :] * np.ones_like(nn)
    sel = np.floor(nn + 0.5) == 5
    assert kk[4] == pytest.approx(2 * np.pi / L * np.sum(w[sel] * nn[sel]) / np.sum(w[sel]), rel=1e-12)


def test_power_spectrum_matches_numpy():
    import ifdm_code_amd as F
    n = 32
    cfg = F.Config.from_cli(2.5e-24, 0.0, 5.0, n)
    psi = make_psi(n, 9, cfg.mfdm_solar)
    with F.Solver(cfg) as s:
        s.set_psi(psi)
        kk, pk, nm = s.power_spectrum(n // 2)
    rho = np.abs(psi) ** 2 * cfg.mfdm_solar
    delta = rho / rho.mean() - 1
    dk = np.fft.rfftn(delta)
    f = np.fft.fftfreq(n) * n
    nz = np.arange(n // 2 + 1)
    nn = np.sqrt(f[:, None, None] ** 2 + f[None, :, None] ** 2 + nz[None, None, :] ** 2)
    w = np.where((nz == 0) | (nz == n // 2), 1.0, 2.0)[None, None, :] * np.ones_like(nn)
    b = np.floor(nn + 0.5).astype(int)
    L = 2 * cfg.xmax
    for bb in range(1, n // 2 + 1):
        sel = b == bb
        ref = L ** 3 * np.sum(w[sel] * np.abs(dk[sel]) ** 2) / n ** 6 / np.sum(w[sel])
        assert pk[bb - 1] == pytest.approx(ref, rel=1e-10)
        assert nm[bb - 1] == np.sum(w[sel])
