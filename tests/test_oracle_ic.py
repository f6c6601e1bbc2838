"""Pins the oracle's initial-condition path (CPU): CIC mass conservation and
known weights, periodic wrap, the reference's float velocity grids, the
phase of a single velocity mode, and the header-derived mean density."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from ifdm_code_amd import icgen


def _rp(n, xmax_in=5.0):
    return O.RunParams(n=n, xmax_in=xmax_in)


def test_cic_mass_conservation_and_uniform_lattice():
    n = 16
    pos, vel, pm, a = icgen.zeldovich(n, 10.0, seed=3)
    box = 10.0e3
    rp = _rp(n)
    pmass = pm * 1e10
    psi, theta, sk = O.ic_from_particles(rp, [(pos, vel)], box, pmass, a)
    assert sk == 0
    # density grid: total deposited mass / cell volume = N^3 pmass / Lg^3
    lg = box / n
    rho = (np.abs(psi) ** 2) * rp.mfdm_solar / (1.225 ** 2 * O.h_0 ** 2 * 1e9)
    assert rho.sum() == pytest.approx(n ** 3 * pmass / lg ** 3, rel=1e-12)
    # an undisplaced lattice deposits exactly one particle mass per cell
    i = (np.arange(n) + 0.5) * lg
    q = np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    psi0, _, _ = O.ic_from_particles(rp, [(q, np.zeros_like(q))], box, pmass, a)
    rho0 = (np.abs(psi0) ** 2) * rp.mfdm_solar / (1.225 ** 2 * O.h_0 ** 2 * 1e9)
    assert np.allclose(rho0, pmass / lg ** 3, rtol=1e-12)
    assert np.abs(np.angle(psi0)).max() == 0.0


def test_cic_weights_and_periodic_wrap():
    n = 8
    box = 10.0e3
    lg = box / n
    rp = _rp(n)
    pmass = 1e9
    # one particle 30% of a cell beyond the centre of cell (2, 7, 0) in x,
    # at the lower edge (wraps to cell N-1) in y, exactly on a centre in z
    p = np.array([[2.8 * lg, 0.1 * lg, 0.5 * lg]], np.float32)
    psi, _, _ = O.ic_from_particles(rp, [(p, np.zeros_like(p))], box, pmass, 0.02)
    rho = (np.abs(psi) ** 2) * rp.mfdm_solar / (1.225 ** 2 * O.h_0 ** 2 * 1e9) * lg ** 3 / pmass
    xp, yp = float(p[0, 0]), float(np.float32(0.1 * lg) + np.float32(box))
    wx1 = abs(xp - (2.5 * lg)) / lg
    wy1 = abs(yp - (7.5 * lg)) / lg
    expect = {(2, 7, 0): (1 - wx1) * (1 - wy1), (2, 0, 0): (1 - wx1) * wy1,
              (3, 7, 0): wx1 * (1 - wy1), (3, 0, 0): wx1 * wy1}
    got = {tuple(int(v) for v in idx): rho[tuple(idx)] for idx in np.argwhere(rho > 1e-14)}
    assert set(got) == set(expect)
    for k, v in expect.items():
        assert got[k] == pytest.approx(v, rel=1e-9)
    assert sum(got.values()) == pytest.approx(1.0, rel=1e-12)


def test_single_velocity_mode_phase():
    """Particles on the cell centres with vx = V sin(2 pi x / L): the CIC
    grids hold the particle values exactly, the reference's central
    difference gives div_i = V u sin(k dx) cos(k x_i) / dx, and for a
    single Fourier mode theta = -a^{3/2} div / k^2.  Box 8192 kpc/h so the
    cell centres are exact in float; the velocities themselves are float,
    which puts ~1e-7 relative noise into other modes (the tolerance)."""
    n = 16
    box = 8192.0
    lg = box / n
    rp = _rp(n, xmax_in=4.096)
    i = (np.arange(n) + 0.5) * lg
    q = np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    V = 40.0
    vel = np.zeros_like(q)
    vel[:, 0] = (V * np.sin(2 * np.pi * q[:, 0].astype(np.float64) / box)).astype(np.float32)
    a = 0.02
    psi, theta, _ = O.ic_from_particles(rp, [(q, vel)], box, 1e9, a)
    unit = 1e3 / O.m2Mpc * O.s2Myr
    vx = vel[:, 0].astype(np.float64).reshape(n, n, n)[:, 0, 0]
    div = 0.5 * (np.roll(vx, -1) - np.roll(vx, 1)) * unit / rp.dx
    kx = 2 * np.pi / (2 * rp.xmax)
    expect = -a ** 1.5 * div / kx ** 2
    assert np.allclose(theta.real[:, 0, 0], expect, rtol=0, atol=2e-7 * np.abs(expect).max())
    assert np.allclose(theta.real, theta.real[:, :1, :1], rtol=0, atol=1e-13 * np.abs(expect).max())
    assert np.abs(theta.imag).max() < 1e-12 * np.abs(expect).max()
    cph = rp.mfdm / O.c_light ** 2 / O.hbar
    assert np.allclose(np.angle(psi[:, 3, 5]), np.angle(np.exp(1j * cph * theta[:, 3, 5].real)), atol=1e-12)


def test_reference_float_velocity_grids_differ_at_float_roundoff():
    n = 16
    pos, vel, pm, a = icgen.zeldovich(n, 10.0, seed=9)
    rp = _rp(n)
    p64, t64, _ = O.ic_from_particles(rp, [(pos, vel)], 10.0e3, pm * 1e10, a, f32=False)
    p32, t32, _ = O.ic_from_particles(rp, [(pos, vel)], 10.0e3, pm * 1e10, a, f32=True)
    d = np.linalg.norm(t32 - t64) / np.linalg.norm(t64)
    assert 1e-9 < d < 1e-5


def float_order_case(order):
    """Three particles on the centre of cell (3, 3, 3) of an 8^3 mesh with
    cell size 1 kpc/h (weight 1 there, 0 elsewhere, Lg^3 = 1): x-velocities
    2^25, 1 and -2^25 visited in `order`.  Returns (pos, vel, box, pmass)."""
    v = {"A": 2.0 ** 25, "B": 1.0, "C": -(2.0 ** 25)}
    pos = np.full((3, 3), 3.5, np.float32)
    vel = np.zeros((3, 3), np.float32)
    vel[:, 0] = [v[c] for c in order]
    return pos, vel, 8.0, 2.0


@pytest.mark.parametrize("order,f32,want", [("ABC", True, 0.0), ("ACB", True, 1.0 / 3), ("ABC", False, 1.0 / 3),
                                            ("ACB", False, 1.0 / 3)])
def test_float_grids_accumulate_in_particle_order(order, f32, want):
    """The reference adds each contribution into a float grid cell
    (vxg[c] += ..., std::vector<float>, src/gadget_IC.h:112-150,583): in
    float, 2^25 + 1 rounds back to 2^25 (the float spacing there is 4), so
    visiting 2^25, 1, -2^25 leaves 0 where 2^25, -2^25, 1 leaves 1; FP64
    grids give 1 either way.  The normalisation then multiplies by
    pmass / rho = 1/3 (:645-651) and rounds to float."""
    pos, vel, box, pmass = float_order_case(order)
    rp = O.RunParams(n=8, xmax_in=5.0)
    _psi, _th, sk, (rho, vx, vy, vz) = O.ic_from_particles(rp, [(pos, vel)], box, pmass, 1 / 51, f32=f32,
                                                           grids=True)
    assert sk == 0
    assert rho[3, 3, 3] == 3 * pmass and np.count_nonzero(rho) == 1
    expect = float(np.float32(want)) if f32 else want
    assert vx[3, 3, 3] == expect
    assert np.count_nonzero(vx) == (1 if want else 0)


def test_ref_p1_ghosts_double_deposit():
    """At one rank the reference deposits the particles beyond the last cell
    centre twice (ghost copy + owner)."""
    n = 8
    box = 10.0e3
    lg = box / n
    rp = _rp(n)
    p = np.array([[1.0 * lg, 2.5 * lg, 3.5 * lg], [7.8 * lg, 2.5 * lg, 3.5 * lg]], np.float32)  # x-sorted
    psi1, _, _ = O.ic_from_particles(rp, [(p, np.zeros_like(p))], box, 1e9, 0.02, ref_p1=False)
    psi2, _, _ = O.ic_from_particles(rp, [(p, np.zeros_like(p))], box, 1e9, 0.02, ref_p1=True)
    m1 = (np.abs(psi1) ** 2).sum()
    m2 = (np.abs(psi2) ** 2).sum()
    assert m2 / m1 == pytest.approx(1.5, rel=1e-12)


def test_header_rhoc_matches_reference_constant():
    n = 16
    _, _, pm, _ = icgen.zeldovich(n, 10.0)
    assert O.header_rhoc(4, n ** 3 // 4, pm * 1e10, 10.0e3) == pytest.approx(O.rhoc_const, rel=1e-12)


def test_gadget_roundtrip(tmp_path):
    n = 8
    pos, vel, pm, a = icgen.zeldovich(n, 10.0, seed=5)
    chunks = icgen.write_gadget(str(tmp_path / "ic"), pos, vel, pm, a, 10.0e3)
    back, hdr = icgen.read_gadget(str(tmp_path / "ic"))
    assert hdr["npart"] == n ** 3 // 4 and hdr["box"] == 10.0e3 and hdr["a"] == a
    for (p1, v1), (p2, v2) in zip(chunks, back):
        assert np.array_equal(p1, p2) and np.array_equal(v1, v2)
        assert np.all(np.diff(p2[:, 0]) >= 0)
