"""GPU: initial conditions, snapshots and the command-line driver against
the oracle and the reference's file formats."""
import os
import subprocess
import sys

import numpy as np
import pytest

from tests.fields import rel_l2

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _say(*a):
    print(*a, flush=True)


def _mods():
    import ifdm_code_amd as F
    from ifdm_code_amd import icgen
    from oracle import oracle as O
    return F, icgen, O


def _sorted_chunks(icgen, n, seed, box_mpc_h=10.0):
    pos, vel, pm, a = icgen.zeldovich(n, box_mpc_h, seed=seed)
    order = np.argsort(pos[:, 0], kind="stable")
    pos, vel = pos[order], vel[order]
    per = n ** 3 // 4
    return [(pos[c * per:(c + 1) * per], vel[c * per:(c + 1) * per]) for c in range(4)], pm, a


@pytest.mark.parametrize("n,ghosts", [(16, False), (32, False), (32, True), (64, True)])
def test_ic_matches_reference_float_deposit(n, ghosts):
    """The default deposit reproduces the reference's arithmetic: the
    density accumulated in FP64 and the velocity grids in float, each cell's
    contributions in the reference's particle order, the divergence in float
    (src/gadget_IC.h:112-150,170,583,645-651).  The mass-weighted float vx
    grid equals the oracle's bit for bit; theta and psi_0 differ only by the
    FFT round-off of the phase solve."""
    F, icgen, O = _mods()
    _say(f"[ic reference order] n={n} ghosts={ghosts}")
    chunks, pm, a = _sorted_chunks(icgen, n, n)
    rp = O.RunParams(n=n, xmax_in=5.0)
    psi_o, th_o, _, (rho_o, vx_o, vy_o, vz_o) = O.ic_from_particles(rp, chunks, 10.0e3, pm * 1e10, a,
                                                                    ref_p1=ghosts, f32=True, grids=True)
    cfg = F.Config.from_cli(2.5e-24, 0.0, 5.0, n)
    with F.Solver(cfg) as s:
        th, skipped, vx = s.ic_from_particles(chunks, 10.0e3, pm * 1e10, a, ref_ghosts=ghosts, want_theta=True,
                                             want_vx=True)
        psi = s.get_psi()
    assert skipped == 0
    assert np.array_equal(vx, vx_o.astype(np.float32))
    e_th, e_psi = rel_l2(th, th_o), rel_l2(psi, psi_o)
    _say(f"   theta {e_th:.3e} psi {e_psi:.3e}")
    assert e_th < 1e-12
    assert e_psi < 1e-12


@pytest.mark.parametrize("n,ghosts", [(16, False), (32, True)])
def test_ic_fp64_atomic_deposit_matches_oracle(n, ghosts):
    """The FP64-atomic deposit (ic_deposit = 1) against the oracle's FP64
    grids: unordered sums, round-off agreement."""
    F, icgen, O = _mods()
    _say(f"[ic atomic] n={n} ghosts={ghosts}")
    chunks, pm, a = _sorted_chunks(icgen, n, n)
    rp = O.RunParams(n=n, xmax_in=5.0)
    psi_o, th_o, _ = O.ic_from_particles(rp, chunks, 10.0e3, pm * 1e10, a, ref_p1=ghosts)
    cfg = F.Config.from_cli(2.5e-24, 0.0, 5.0, n)
    with F.Solver(cfg) as s:
        s.set_option("ic_deposit", 1)
        th, skipped = s.ic_from_particles(chunks, 10.0e3, pm * 1e10, a, ref_ghosts=ghosts, want_theta=True)
        psi = s.get_psi()
    assert skipped == 0
    assert rel_l2(th, th_o) < 1e-11
    assert rel_l2(psi, psi_o) < 1e-12


def test_ic_reference_deposit_is_deterministic():
    """Two runs of the default deposit give psi_0 bit for bit (the order of
    every cell's additions is fixed), also with ghosts and several chunks;
    particles outside the box are counted and left out, as by the oracle."""
    F, icgen, O = _mods()
    n = 32
    chunks, pm, a = _sorted_chunks(icgen, n, 77)
    # three particles the reference cannot place (x below the wrapped half
    # cell, x and y beyond the last cell): skipped; a y below the box wraps
    # (src/gadget_IC.h:67-80) and is deposited
    p0, v0 = chunks[0]
    p0 = p0.copy()
    p0[5, 0] = -500.0
    p0[9, 0] = 10.2e3
    p0[12, 1] = 10.2e3
    p0[20, 1] = -400.0
    chunks[0] = (p0, v0)
    cfg = F.Config.from_cli(2.5e-24, 0.0, 5.0, n)
    outs = []
    for _ in range(2):
        with F.Solver(cfg) as s:
            _th, skipped, vx = s.ic_from_particles(chunks, 10.0e3, pm * 1e10, a, ref_ghosts=True, want_vx=True)
            outs.append((s.get_psi(), vx, skipped))
    assert outs[0][2] == 3 and outs[1][2] == 3
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    rp = O.RunParams(n=n, xmax_in=5.0)
    psi_o, _th, sk_o, grids = O.ic_from_particles(rp, chunks, 10.0e3, pm * 1e10, a, ref_p1=True, f32=True,
                                                  grids=True)
    assert sk_o == 3
    assert np.array_equal(outs[0][1], grids[1].astype(np.float32))
    assert rel_l2(outs[0][0], psi_o) < 1e-12


def test_ic_mixed_precision():
    F, icgen, O = _mods()
    n = 32
    _say(f"[ic mixed] n={n}")
    pos, vel, pm, a = icgen.zeldovich(n, 10.0, seed=2)
    rp = O.RunParams(n=n)
    psi_o, _, _ = O.ic_from_particles(rp, [(pos, vel)], 10.0e3, pm * 1e10, a, f32=True)
    cfg = F.Config.from_cli(2.5e-24, 0.0, 5.0, n, precision=F.MIXED)
    with F.Solver(cfg) as s:
        s.ic_from_particles([(pos, vel)], 10.0e3, pm * 1e10, a)
        psi = s.get_psi()
    assert rel_l2(psi, psi_o) < 1e-7


def test_read_gadget(tmp_path):
    F, icgen, O = _mods()
    n = 16
    _say(f"[read_gadget] n={n}")
    pos, vel, pm, a = icgen.zeldovich(n, 10.0, seed=4)
    chunks = icgen.write_gadget(str(tmp_path / "ic" / "sorted"), pos, vel, pm, a, 10.0e3)
    rp = O.RunParams(n=n)
    psi_o, _, _, grids = O.ic_from_particles(rp, chunks, 10.0e3, pm * 1e10, a, f32=True, grids=True)
    cfg = F.Config.from_cli(2.5e-24, 0.0, 5.0, n)
    vout = str(tmp_path / "voutput.dat")
    with F.Solver(cfg) as s:
        info = s.read_gadget(str(tmp_path / "ic" / "sorted"), 4, voutput=vout)
        psi = s.get_psi()
    assert info["npart"] == n ** 3 and info["skipped"] == 0 and info["a"] == pytest.approx(a)
    assert info["rhoc"] == pytest.approx(O.header_rhoc(4, n ** 3 // 4, pm * 1e10, 10.0e3), rel=1e-14)
    assert rel_l2(psi, psi_o) < 1e-12
    # voutput.dat: the float vx grid after the velocity normalisation, in
    # the reference layout (write_binary_vg, src/write3D.h:24-53)
    assert os.path.getsize(vout) == 4 * n ** 3
    assert np.array_equal(np.fromfile(vout, np.float32), grids[1].astype(np.float32).ravel())


def test_read_gadget_errors(tmp_path):
    F, icgen, O = _mods()
    n = 8
    pos, vel, pm, a = icgen.zeldovich(n, 10.0, seed=4)
    icgen.write_gadget(str(tmp_path / "ic"), pos, vel, pm, a, 10.0e3)
    # box mismatch: the IC is 10 Mpc/h, the run asks for xmax 4 (8 Mpc/h)
    with F.Solver(F.Config.from_cli(2.5e-24, 0.0, 4.0, n)) as s:
        with pytest.raises(F.IfdmError, match="Boxsize"):
            s.read_gadget(str(tmp_path / "ic"), 4)
    with F.Solver(F.Config.from_cli(2.5e-24, 0.0, 5.0, n)) as s:
        with pytest.raises(F.IfdmError, match="cannot read IC file"):
            s.read_gadget(str(tmp_path / "missing"), 4)
        # a truncated chunk
        with open(tmp_path / "ic.3", "r+b") as f:
            f.truncate(1000)
        with pytest.raises(F.IfdmError, match="truncated"):
            s.read_gadget(str(tmp_path / "ic"), 4)


@pytest.mark.parametrize("prec", [0, 1])
def test_snapshot_layout_and_rerun(tmp_path, prec):
    """output<t>.dat = [mfdm][g][a][x[N]][psi N^3 complex128] (write3D.h),
    read back by the rerun path; a restart continues bit-identically."""
    F, icgen, O = _mods()
    n = 16
    _say(f"[snapshot] prec={prec}")
    pos, vel, pm, a = icgen.zeldovich(n, 10.0, seed=6)
    cfg = F.Config.from_cli(2.5e-24, 1e-60, 5.0, n, precision=prec)
    path = str(tmp_path / "output7.dat")
    with F.Solver(cfg) as s:
        s.ic_from_particles([(pos, vel)], 10.0e3, pm * 1e10, a)
        a1, sched = s.simulate_native(a, 3)
        s.write_snapshot(path, a1, 2.5e-24, 1e-60)
        psi = s.get_psi()
        a2, _ = s.simulate_native(a1, 3)
        cont = s.get_psi()
    raw = np.fromfile(path, np.float64)
    assert raw.size == 3 + n + 2 * n ** 3
    assert raw[0] == 2.5e-24 and raw[1] == 1e-60 and raw[2] == a1
    xmax = cfg.xmax
    assert np.allclose(raw[3:3 + n], xmax / n - xmax + np.arange(n) * (2 * xmax / n), rtol=0, atol=1e-15)
    assert np.array_equal(raw[3 + n:].view(np.complex128).reshape(n, n, n), psi)
    with F.Solver(cfg) as s2:
        ar, m, g = s2.read_snapshot(path)
        assert (ar, m, g) == (a1, 2.5e-24, 1e-60)
        assert np.array_equal(s2.get_psi(), psi) if prec == 0 else rel_l2(s2.get_psi(), psi) < 1e-7
        a3, _ = s2.simulate_native(ar, 3)
        again = s2.get_psi()
    if prec == 0:
        assert a3 == a2
        assert np.array_equal(again, cont)
    else:
        # the mixed Poisson source subtracts the FP64 mean density, which the
        # restart recomputes on the host: FP32-level differences in phi and dt
        assert a3 == pytest.approx(a2, rel=1e-6)
        assert rel_l2(again, cont) < 1e-5


def test_readsnap_reads_our_snapshot(tmp_path):
    """The reference's own reader (src/readsnap.cpp, built into oracle/_ref)
    turns a library-written snapshot into the expected ASCII z-slice."""
    exe = os.path.join(ROOT, "oracle", "_ref", "readsnap")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/readsnap not built (needs the reference sources at build time)")
    F, icgen, O = _mods()
    n = 8
    pos, vel, pm, a = icgen.zeldovich(n, 10.0, seed=8)
    cfg = F.Config.from_cli(2.5e-24, 0.0, 5.0, n)
    d = str(tmp_path) + "/"
    with F.Solver(cfg) as s:
        s.ic_from_particles([(pos, vel)], 10.0e3, pm * 1e10, a)
        s.write_snapshot(d + "output5.dat", a, 2.5e-24, 0.0)
        psi = s.get_psi()
    r = subprocess.run([exe, str(n), "3", "5", d, d], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    rows = np.loadtxt(d + "output5.dat.z3")
    assert rows.shape == (n * n, 4)
    ms = 2.5e-24 * O.kg2eV / O.kg2solar
    x = cfg.xmax / n - cfg.xmax + np.arange(n) * (2 * cfg.xmax / n)
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    expect = np.stack([x[i].ravel(), x[j].ravel(), (np.abs(psi[:, :, 3]) ** 2 * ms).ravel(),
                       np.angle(psi[:, :, 3]).ravel()], axis=1)
    assert np.allclose(rows, expect, rtol=2e-5, atol=1e-5 * np.abs(expect).max(axis=0))


def test_cli_matches_library_and_restarts(tmp_path):
    """The 10-parameter driver on a Gadget IC, then a rerun from one of its
    snapshots, against the library driven step by step from Python."""
    F, icgen, O = _mods()
    n = 16
    _say(f"[cli] n={n}")
    exe = os.path.join(ROOT, "ifdm_code_amd", "ifdm")
    pos, vel, pm, a0 = icgen.zeldovich(n, 10.0, seed=12)
    icgen.write_gadget(str(tmp_path / "ic" / "sorted"), pos, vel, pm, a0, 10.0e3)
    out = str(tmp_path / "out") + "/"
    os.makedirs(out)
    args = [exe, "2.5e-24", "0", "5.", str(n), "3", "7", "0", out, str(tmp_path / "ic" / "sorted"), "0",
            "--no-a-triggers"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    snaps = sorted(os.listdir(out))
    assert "output1.dat" in snaps and "output4.dat" in snaps and "output7.dat" in snaps and "voutput.dat" in snaps
    # the same run through the Python binding: snapshot t holds psi before step t-1
    cfg = F.Config.from_cli(2.5e-24, 0.0, 5.0, n)
    with F.Solver(cfg) as s:
        info = s.read_gadget(str(tmp_path / "ic" / "sorted"), 4)
        st = s.poisson()
        a = info["a"]
        trig = F.OutputTriggers(0, False)
        for tn in range(7):
            dt, a, _ = F.timestep_criteria(cfg, a, st)
            t, _stop = trig(a, tn, 3, 7)
            if t is not None:
                raw = np.fromfile(out + f"output{t}.dat", np.float64)
                # the deposit fixes every cell's order: the two runs agree bit for bit
                assert raw[2] == a
                assert np.array_equal(raw[3 + n:].view(np.complex128).reshape(n, n, n), s.get_psi())
            st = s.step(*dt)
    # rerun from output4.dat (ti = 4) for the remaining steps
    out2 = str(tmp_path / "out2") + "/"
    os.makedirs(out2)
    args = [exe, "2.5e-24", "0", "5.", str(n), "3", "7", "4", out2, out, "1", "--no-a-triggers"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert os.path.exists(out2 + "output7.dat")


def test_snapshot_bytes_match_the_golden_layout(tmp_path):
    """ifdm_write_snapshot writes exactly the bytes of the layout that the
    reference's reader was shown to read (tests/golden/readsnap_n8_z3.txt),
    and ifdm_read_snapshot restores psi bit for bit."""
    from tests.fields import GOLDEN_SNAP, make_psi, snapshot_bytes
    F, _icgen, O = _mods()
    g = GOLDEN_SNAP
    _say(f"[snapshot golden] n={g['n']}")
    cfg = F.Config.from_cli(g["mfdm_ev"], g["g_ev_cm3"], g["xmax_mpc_h"], g["n"])
    psi = make_psi(g["n"], g["seed"], cfg.mfdm_solar)
    path = str(tmp_path / "output1.dat")
    with F.Solver(cfg) as s:
        s.set_psi(psi)
        s.write_snapshot(path, g["a"], g["mfdm_ev"], g["g_ev_cm3"])
        s.set_psi(np.zeros_like(psi))
        a, m, gg = s.read_snapshot(path)
        back = s.get_psi()
    assert open(path, "rb").read() == snapshot_bytes(psi, cfg.xmax, g["a"], g["mfdm_ev"], g["g_ev_cm3"])
    assert (a, m, gg) == (g["a"], g["mfdm_ev"], g["g_ev_cm3"])
    assert np.array_equal(back, psi)


def test_cli_reference_run_64(tmp_path):
    """The reference's run with its ten parameters, `2.5e-24 0 5 64 100 200 0
    out/ ic 0` (64^3 Zel'dovich IC, box 10 Mpc/h, z = 50, seed 1234, in the
    4-chunk sorted Gadget layout): the driver runs the reference's trigger
    chain (src/CosmoBEC_MPI_stan.cpp:238-284) to its end of run at a > 0.5
    and must write the same snapshots as the oracle's run
    (tests/golden/cli_64_oracle.npz): the same numbers, a to 1e-12, psi on
    the fixture's sample cells to 1e-10, the norm to 1e-12, and voutput.dat
    byte for byte."""
    import hashlib
    F, icgen, O = _mods()
    fx = np.load(os.path.join(ROOT, "tests", "golden", "cli_64_oracle.npz"))
    n = 64
    _say(f"[cli reference run] n={n}, {int(fx['steps'])} steps, {len(fx['t'])} snapshots")
    exe = os.path.join(ROOT, "ifdm_code_amd", "ifdm")
    pos, vel, pm, a0 = icgen.zeldovich(n, 10.0, z=50.0, seed=1234)
    icgen.write_gadget(str(tmp_path / "ic" / "sorted"), pos, vel, pm, a0, 10.0e3)
    out = str(tmp_path / "out") + "/"
    os.makedirs(out)
    args = [exe, "2.5e-24", "0", "5", str(n), "100", "200", "0", out, str(tmp_path / "ic" / "sorted"), "0"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    written = sorted(int(f[6:-4]) for f in os.listdir(out) if f.startswith("output") and f.endswith(".dat"))
    assert written == sorted(int(t) for t in fx["t"])
    idx = fx["idx"]
    worst = 0.0
    for k, t in enumerate(fx["t"]):
        raw = np.fromfile(out + f"output{int(t)}.dat", np.float64)
        assert raw.size == 3 + n + 2 * n ** 3
        assert raw[2] == pytest.approx(float(fx["a"][k]), rel=1e-12)
        psi = raw[3 + n:].view(np.complex128)
        err = rel_l2(psi[idx], fx["psi"][k])
        worst = max(worst, err)
        assert err < 1e-10, (int(t), err)
        assert abs(float(np.sum(np.abs(psi) ** 2)) / float(fx["norm"][k]) - 1) < 1e-12
    _say(f"   worst sampled psi rel L2 over the snapshots: {worst:.3e}")
    with open(out + "voutput.dat", "rb") as f:
        assert hashlib.sha256(f.read()).hexdigest() == str(fx["voutput_sha256"])


@pytest.mark.parametrize("order,mode,want", [("ABC", 0, 0.0), ("ACB", 0, 1.0 / 3), ("ABC", 1, 1.0 / 3),
                                             ("ACB", 1, 1.0 / 3)])
def test_deposit_float_rounding_in_particle_order(order, mode, want):
    """Three particles on one cell centre with x-velocities 2^25, 1, -2^25
    (tests/test_oracle_ic.py: float_order_case): the default deposit rounds
    the float grid cell after each addition in the reference's particle
    order (2^25, 1, -2^25 leaves 0; 2^25, -2^25, 1 leaves 1, then x 1/3 from
    the normalisation); the FP64-atomic deposit gives 1/3 either way."""
    from tests.test_oracle_ic import float_order_case
    F, _icgen, _O = _mods()
    pos, vel, box, pmass = float_order_case(order)
    with F.Solver(F.Config.from_cli(2.5e-24, 0.0, 5.0, 8)) as s:
        s.set_option("ic_deposit", mode)
        _th, sk, vx = s.ic_from_particles([(pos, vel)], box, pmass, 1 / 51, want_vx=True)
    assert sk == 0
    assert vx[3, 3, 3] == np.float32(want)
    assert np.count_nonzero(vx) == (1 if want else 0)
