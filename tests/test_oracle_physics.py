"""Pins the oracle's physics against closed forms of the comoving
Schrodinger-Poisson equations the reference integrates
(i hbar dpsi/dt = -hbar^2/(2 m a^2) lap psi + m phi/a psi + g|psi|^2/a^3 psi,
with the three time integrals dKE, dPE, dQI).  CPU only."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.fields import make_psi, rel_l2


def _grid_index(n):
    return np.meshgrid(*[np.arange(n)] * 3, indexing="ij")


@pytest.mark.parametrize("n,m", [(16, (1, 0, 0)), (32, (2, 1, 3)), (32, (0, 0, 5))])
def test_poisson_cosine_closed_form(n, m):
    """rho = rho0 (1 + eps cos(k0.x)) -> phi = -4 pi G rho0 eps cos(k0.x)/|k0|^2."""
    rp = O.RunParams(n=n)
    i, j, k = _grid_index(n)
    arg = 2 * np.pi * (m[0] * i + m[1] * j + m[2] * k) / n
    eps = 0.03
    rho = rp.rhoc * (1 + eps * np.cos(arg))
    st = O.State(rp, np.sqrt(rho / rp.mfdm_solar))
    st.poisson()
    L = 2 * rp.xmax
    k0sq = sum((2 * np.pi * mm / L) ** 2 for mm in m)
    exact = -4 * np.pi * O.bigG * rp.rhoc * eps * np.cos(arg) / k0sq
    assert rel_l2(st.phi, exact) < 1e-12


def test_uniform_field_self_interaction_phase():
    """Uniform psi: phi = 0 and no drift, so one step multiplies psi by
    exp(-i g |psi|^2 dQI / hbar) (two half-kicks)."""
    n = 16
    rp = O.RunParams(n=n, g_in=4e-55)
    A = math.sqrt(rp.rhoc / rp.mfdm_solar) * np.exp(0.3j)
    st = O.State(rp, np.full((n, n, n), A))
    st.poisson()
    assert np.abs(st.phi).max() < 1e-12 * (4 * np.pi * O.bigG * rp.rhoc * rp.xmax ** 2)
    dKE, dPE, dQI = 5e3, 1e-3, 0.37
    st.split_op(dKE, dPE, dQI)
    expect = A * np.exp(-1j * rp.g / O.hbar * abs(A) ** 2 * dQI)
    assert np.abs(st.psi - expect).max() / abs(A) < 1e-12


def test_potential_phase_closed_form():
    """Cosine density, no drift (dKE = 0): both half-kicks see the same
    analytic phi, so psi' = psi exp(-i m phi dPE / (c^2 hbar))."""
    n = 32
    rp = O.RunParams(n=n)
    i, j, k = _grid_index(n)
    arg = 2 * np.pi * (i + 2 * k) / n
    eps = 0.05
    rho = rp.rhoc * (1 + eps * np.cos(arg))
    psi = np.sqrt(rho / rp.mfdm_solar) * np.exp(0.1j * np.sin(arg))
    st = O.State(rp, psi)
    st.poisson()
    L = 2 * rp.xmax
    phi = -4 * np.pi * O.bigG * rp.rhoc * eps * np.cos(arg) / ((2 * np.pi / L) ** 2 * 5)
    dPE = 3.0e3
    st.split_op(0.0, dPE, 0.0)
    mt = rp.mfdm / O.c_light ** 2
    expect = psi * np.exp(-1j * mt / O.hbar * phi * dPE)
    assert rel_l2(st.psi, expect) < 1e-10


def test_drift_plane_waves():
    """sum_m A_m exp(i k_m.x) evolves as A_m exp(-i hbar |k_m|^2 dKE / (2 m))."""
    n = 32
    rp = O.RunParams(n=n)
    i, j, k = _grid_index(n)
    modes = [((1, 0, 0), 1.0), ((3, -2, 5), 0.5 + 0.2j), ((-16, 7, 0), 0.3j), ((0, 0, 15), -0.25)]
    L = 2 * rp.xmax
    psi = np.zeros((n, n, n), complex)
    out = np.zeros_like(psi)
    dKE = 2.0e4
    mt = rp.mfdm / O.c_light ** 2
    for (mx, my, mz), amp in modes:
        wave = np.exp(2j * np.pi * (mx * i + my * j + mz * k) / n)
        ksq = (2 * np.pi / L) ** 2 * (mx * mx + my * my + mz * mz)
        psi += amp * wave
        out += amp * wave * np.exp(-1j * O.hbar / (2 * mt) * ksq * dKE)
    st = O.State(rp, psi)
    st.drift(dKE)
    assert rel_l2(st.psi, out) < 1e-13


def test_norm_conservation_and_time_reversal():
    n = 16
    rp = O.RunParams(n=n, g_in=4e-55)
    psi = make_psi(n, 4, rp.mfdm_solar)
    st = O.State(rp, psi)
    st.poisson()
    m = st.maxima()
    dt = O.timestep_criteria(rp, 1 / 51, m["max_phi"], m["max_rho"])[:3]
    n0 = np.sum(np.abs(psi) ** 2)
    for _ in range(3):
        st.split_op(*dt)
    assert abs(np.sum(np.abs(st.psi) ** 2) / n0 - 1) < 1e-13
    # reverse: conjugate, same steps, conjugate
    st2 = O.State(rp, np.conj(st.psi))
    st2.poisson()
    for _ in range(3):
        st2.split_op(*dt)
    assert rel_l2(np.conj(st2.psi), psi) < 1e-12


def test_maxima_ranges():
    n = 16
    rp = O.RunParams(n=n)
    st = O.State(rp, make_psi(n, 8, rp.mfdm_solar))
    st.poisson()
    full = st.maxima(False)
    ref = st.maxima(True)
    flat = st.phi_pad.ravel()
    assert ref["max_phi"] == np.abs(flat[: n ** 3]).max()
    assert full["max_phi"] == np.abs(st.phi).max()
    assert full["max_rho"] == pytest.approx((np.abs(st.psi) ** 2).max(), rel=1e-15)
    assert full["norm"] == pytest.approx((np.abs(st.psi) ** 2).sum(), rel=1e-13)


def _phase_per_step(rp, a, m, which_expected):
    dKE, dPE, dQI, a1, which = O.timestep_criteria(rp, a, m["max_phi"], m["max_rho"])
    assert which == which_expected
    mt = rp.mfdm / O.c_light ** 2
    if which == 3:
        kmax2 = 3 * (math.pi / rp.dx) ** 2
        ph = 0.5 * O.hbar / mt * kmax2 * dKE
    elif which == 2:
        ph = mt / O.hbar * m["max_phi"] * dPE
    else:
        ph = rp.g / O.hbar * m["max_rho"] * dQI
    return ph, (a1 - a) / a


@pytest.mark.parametrize("case", ["KE", "PE", "QI"])
def test_timestep_gives_2pi_phase(case):
    """The criteria are built so that the limiting phase of one step is 2 pi
    as da -> 0; the integrals over the finite step deviate by O(da/a)."""
    if case == "KE":
        rp, m, w = O.RunParams(n=512), {"max_phi": 1e-30, "max_rho": 1.0}, 3
    elif case == "PE":
        rp, m, w = O.RunParams(n=64), {"max_phi": 5e-7, "max_rho": 1.0}, 2
    else:
        rp, m, w = O.RunParams(n=64, g_in=4e-55), {"max_phi": 1e-30, "max_rho": 2e100}, 4
    ph, rel_da = _phase_per_step(rp, 1 / 51, m, w)
    assert rel_da < 0.05
    assert abs(ph / (2 * math.pi) - 1) < 3 * rel_da


@pytest.mark.parametrize("n_pow", [2, 3, 4])
@pytest.mark.parametrize("a0,a1", [(1 / 51, 1 / 51 * 1.01), (0.02, 0.03), (0.05, 1.0)])
def test_integrals_vs_mpmath(n_pow, a0, a1):
    import mpmath as mp
    mp.mp.dps = 40
    H0, Om = mp.mpf(O.H_0), mp.mpf(O.Omega_m0)
    f = lambda x: 1 / (H0 * mp.sqrt(Om / x ** 3 + (1 - Om)) * x ** n_pow)
    ref = float(mp.quad(f, [a0, a1]))
    assert O.integral(a0, a1, n_pow) == pytest.approx(ref, rel=1e-13)


def test_fft_library_restatement_closed_forms_and_oracle():
    """FastState (the step over scipy.fft, the bench's CPU baseline on a
    library transform) against the same closed forms as the C oracle - the
    single-mode potential and the plane-wave drift - and against the C
    oracle (State) over three adaptive steps with g != 0."""
    n = 32
    rp = O.RunParams(n=n, g_in=4e-55)
    i, j, k = _grid_index(n)
    m = (2, 1, 3)
    arg = 2 * np.pi * (m[0] * i + m[1] * j + m[2] * k) / n
    rho = rp.rhoc * (1 + 0.03 * np.cos(arg))
    fs = O.FastState(rp, np.sqrt(rho / rp.mfdm_solar), workers=4)
    fs.poisson()
    L = 2 * rp.xmax
    k0sq = sum((2 * np.pi * mm / L) ** 2 for mm in m)
    assert rel_l2(fs.phi, -4 * np.pi * O.bigG * rp.rhoc * 0.03 * np.cos(arg) / k0sq) < 1e-12
    wave = np.exp(2j * np.pi * (3 * i - 2 * j + 5 * k) / n)
    ksq = (2 * np.pi / L) ** 2 * 38
    fs.psi = wave.copy()
    fs.drift(2.0e4)
    assert rel_l2(fs.psi, wave * np.exp(-1j * O.hbar / (2 * rp.mfdm / O.c_light ** 2) * ksq * 2.0e4)) < 1e-13
    fs.close()
    psi = make_psi(n, 6, rp.mfdm_solar)
    a, b = O.State(rp, psi), O.FastState(rp, psi, workers=4)
    t = 1.0 / 51
    for _ in range(3):
        a.poisson()
        b.poisson()
        ma, mb = a.maxima(), b.maxima()
        for key in ("max_phi", "max_rho", "norm"):
            assert mb[key] == pytest.approx(ma[key], rel=1e-12)
        dKE, dPE, dQI, t, _w = O.timestep_criteria(rp, t, ma["max_phi"], ma["max_rho"])
        a.split_op(dKE, dPE, dQI)
        b.split_op(dKE, dPE, dQI)
    assert rel_l2(b.psi, a.psi) < 1e-13
    b.close()
