"""Host-side physics of the library (CPU): the Gauss-Kronrod table, the
scale-factor integrals and Timestep_criteria against the oracle."""
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def gk_table():
    src = open(os.path.join(ROOT, "ifdm_code_amd", "csrc", "criteria.h")).read()
    out = {}
    for name in ("xgk", "wgk", "wg"):
        body = re.search(r"#define IFDM_GK21_%s\s*\\\s*\{(.*?)\}" % name.upper(), src, re.S).group(1)
        out[name] = np.array([float(v) for v in re.findall(r"[-+0-9.eE]+", body)])
    return out


def test_gauss_kronrod_exactness():
    """The 21-point Kronrod rule integrates x^k exactly for k <= 31 and the
    embedded 10-point Gauss rule for k <= 19."""
    t = gk_table()
    xk = np.concatenate([t["xgk"][:10], [0.0], -t["xgk"][:10]])
    wk = np.concatenate([t["wgk"][:10], [t["wgk"][10]], t["wgk"][:10]])
    xg = np.concatenate([t["xgk"][1:10:2], -t["xgk"][1:10:2]])
    wgg = np.concatenate([t["wg"], t["wg"]])
    for k in range(32):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(math.fsum(wk * xk ** k) - exact) < 3e-16, k
        if k < 20:
            assert abs(math.fsum(wgg * xg ** k) - exact) < 3e-16, k
    # and not beyond: degree 32 (Kronrod) / 20 (Gauss) leave a residual
    assert abs(np.sum(wk * xk ** 32) - 2.0 / 33) > 1e-12
    assert abs(np.sum(wgg * xg ** 20) - 2.0 / 21) > 1e-8


@pytest.mark.parametrize("power", [2, 3, 4])
@pytest.mark.parametrize("a0,a1", [(1 / 51, 1 / 51 * 1.002), (0.02, 0.031), (0.02, 1.0), (0.3, 0.5)])
def test_scale_integral_vs_mpmath(power, a0, a1):
    import mpmath as mp
    import ifdm_code_amd as F
    from oracle import oracle as O
    mp.mp.dps = 40
    H0, Om = mp.mpf(O.H_0), mp.mpf(O.Omega_m0)
    ref = float(mp.quad(lambda x: 1 / (H0 * mp.sqrt(Om / x ** 3 + (1 - Om)) * x ** power), [a0, a1]))
    assert F.scale_integral(a0, a1, power) == pytest.approx(ref, rel=1e-13)


@pytest.mark.parametrize("n,g_in,maxphi,maxrho", [
    (64, 0.0, 3.7e-10, 1.9e100), (128, 0.0, 5e-7, 1.9e100), (512, 0.0, 3e-10, 2e100),
    (64, 4e-55, 3.7e-10, 1.9e100), (32, 1e-56, 1e-9, 3e100)])
@pytest.mark.parametrize("a", [1 / 51, 0.1, 0.45])
def test_timestep_criteria_matches_oracle(n, g_in, maxphi, maxrho, a):
    import ifdm_code_amd as F
    from oracle import oracle as O
    cfg = F.Config.from_cli(2.5e-24, g_in, 5.0, n)
    rp = O.RunParams(n=n, g_in=g_in)
    assert cfg.xmax == rp.xmax and cfg.g == rp.g
    (dKE, dPE, dQI), a1, w = F.timestep_criteria(cfg, a, {"max_phi": maxphi, "max_rho": maxrho})
    oKE, oPE, oQI, oa1, ow = O.timestep_criteria(rp, a, maxphi, maxrho)
    assert w == ow
    assert a1 == oa1
    for x, y in [(dKE, oKE), (dPE, oPE), (dQI, oQI)]:
        assert x == pytest.approx(y, rel=1e-13)
