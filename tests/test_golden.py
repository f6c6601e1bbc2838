"""Checks against the committed golden fixtures (tests/golden, written by
tests/golden/make_golden.py)."""
import json
import os

import numpy as np
import pytest

from tests.fields import GOLDEN_SNAP, make_psi, rel_l2

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_readsnap_fixture_matches_the_layout():
    """The reference's own reader (src/readsnap.cpp:46-70) read the golden
    snapshot as x_i, x_j, m_s |psi|^2, arg psi at z = 3: the byte layout
    written by tests.fields.snapshot_bytes (and by ifdm_write_snapshot, whose
    bytes the GPU test compares with it) is the reference's."""
    from oracle import oracle as O
    g = GOLDEN_SNAP
    rp = O.RunParams(n=g["n"], mfdm=g["mfdm_ev"], g_in=g["g_ev_cm3"], xmax_in=g["xmax_mpc_h"])
    psi = make_psi(g["n"], g["seed"], rp.mfdm_solar)
    rows = np.loadtxt(os.path.join(GOLDEN, "readsnap_n8_z3.txt"))
    n = g["n"]
    x = np.array([rp.xmax / n - rp.xmax + i * (2.0 * rp.xmax / n) for i in range(n)])
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    sl = psi[:, :, g["zslice"]]
    expect = np.stack([x[i].ravel(), x[j].ravel(), (np.abs(sl) ** 2 * rp.mfdm_solar).ravel(),
                       np.angle(sl).ravel()], axis=1)
    # readsnap prints 6 significant digits
    assert np.allclose(rows, expect, rtol=1e-5, atol=0)


def test_scale_integrals_fixture():
    import ifdm_code_amd as F
    from oracle import oracle as O
    fx = json.load(open(os.path.join(GOLDEN, "scale_integrals.json")))
    assert fx["H_0"] == O.H_0 and fx["Omega_m0"] == O.Omega_m0
    for c in fx["cases"]:
        v = float(c["value"])
        assert F.scale_integral(c["a0"], c["a1"], c["power"]) == pytest.approx(v, rel=1e-13)
        assert O.integral(c["a0"], c["a1"], c["power"]) == pytest.approx(v, rel=1e-10)


def test_oracle_kdk_fixture():
    from oracle import oracle as O
    fx = np.load(os.path.join(GOLDEN, "kdk_16_oracle.npz"))
    rp = O.RunParams(n=16)
    st, used = O.simulate(rp, make_psi(16, 5, rp.mfdm_solar), 1.0 / 51, 3)
    assert np.allclose(np.array([u[:3] for u in used]), fx["sched"], rtol=1e-14, atol=0)
    assert rel_l2(st.psi, fx["psi"]) < 1e-13
