"""GPU: the hand-written adjoint of the KDK step against central finite
differences (eps = 1e-6) along seeded random directions, through both the
library's forward and the CPU oracle's, and through torch.autograd."""
import numpy as np
import pytest

from tests.fields import make_psi, rel_l2

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _setup(n, steps, g_in=0.0, seed=3, tseed=4321):
    import ifdm_code_amd as F
    from oracle import oracle as O
    rp = O.RunParams(n=n, g_in=g_in)
    cfg = F.Config.from_cli(2.5e-24, g_in, 5.0, n)
    psi0 = make_psi(n, seed, cfg.mfdm_solar)
    # frozen schedule: adaptive steps of the reference loop
    _st, used = O.simulate(rp, psi0, 1.0 / 51, steps)
    sched = [u[:3] for u in used]
    tgt_state, _ = O.simulate(rp, make_psi(n, tseed, cfg.mfdm_solar), 1.0 / 51, steps)
    rho_t = np.abs(tgt_state.psi) ** 2
    rbar = float(np.mean(np.abs(psi0) ** 2))
    return F, O, rp, cfg, psi0, sched, rho_t, rbar


def _loss(psi_T, rho_t, rbar):
    return float(np.sum(((np.abs(psi_T) ** 2 - rho_t) / rbar) ** 2))


def _seed(psi_T, rho_t, rbar):
    return 4 * (np.abs(psi_T) ** 2 - rho_t) / rbar ** 2 * psi_T


@pytest.mark.parametrize("n,steps,K,g_in", [(16, 12, 4, 0.0), (32, 20, 5, 4e-55)])
def test_adjoint_matches_finite_differences(n, steps, K, g_in):
    F, O, rp, cfg, psi0, sched, rho_t, rbar = _setup(n, steps, g_in)
    _say(f"[adjoint] n={n} steps={steps} K={K} g={g_in}")
    with F.Solver(cfg) as s:
        def fwd(p):
            s.set_psi(p)
            s.poisson()
            s.run(sched)
            return s.get_psi()
        s.set_psi(psi0)
        s.poisson()
        s.run_checkpointed(sched, K)
        psiT = s.get_psi()
        g0 = s.adjoint(sched, _seed(psiT, rho_t, rbar))
        assert rel_l2(s.get_psi(), psi0) < 1e-11          # psi_0 restored
        rng = np.random.default_rng(77)
        scale = np.sqrt(rbar)
        eps = 1e-6
        for k in range(4):
            d = (rng.standard_normal(psi0.shape) + 1j * rng.standard_normal(psi0.shape)) * scale / np.sqrt(2)
            fd = (_loss(fwd(psi0 + eps * d), rho_t, rbar) - _loss(fwd(psi0 - eps * d), rho_t, rbar)) / (2 * eps)
            ad = float(np.sum((np.conj(g0) * d).real))
            _say(f"   dir {k}: adjoint {ad:.12e} fd {fd:.12e}")
            assert abs(ad - fd) <= 1e-6 * abs(fd)


def test_adjoint_against_oracle_finite_differences():
    """Finite differences of the CPU oracle's forward (the reference's
    unmerged step sequence) check the GPU gradient independently."""
    n, steps = 16, 8
    F, O, rp, cfg, psi0, sched, rho_t, rbar = _setup(n, steps, 0.0, seed=11)
    with F.Solver(cfg) as s:
        s.set_psi(psi0)
        s.poisson()
        s.run_checkpointed(sched, 4)
        g0 = s.adjoint(sched, _seed(s.get_psi(), rho_t, rbar))

    def fwd_oracle(p):
        st = O.State(rp, p)
        for dt in sched:
            st.poisson()
            st.split_op(*dt)
        return st.psi
    rng = np.random.default_rng(5)
    eps = 1e-6
    for k in range(3):
        d = (rng.standard_normal(psi0.shape) + 1j * rng.standard_normal(psi0.shape)) * np.sqrt(rbar / 2)
        fd = (_loss(fwd_oracle(psi0 + eps * d), rho_t, rbar) - _loss(fwd_oracle(psi0 - eps * d), rho_t, rbar)) / (2 * eps)
        ad = float(np.sum((np.conj(g0) * d).real))
        assert abs(ad - fd) <= 1e-6 * abs(fd)


def test_checkpoints_serve_one_matching_sweep():
    """ifdm_run_checkpointed solves Poisson for the entry state itself, and
    its checkpoints serve exactly one reverse sweep over the same schedule:
    a second sweep, a sweep over another schedule, or a sweep after psi was
    replaced is refused (IFDM_E_STATE) instead of returning the gradient of
    another function."""
    n, steps = 16, 6
    F, O, rp, cfg, psi0, sched, rho_t, rbar = _setup(n, steps, 0.0, seed=31)
    with F.Solver(cfg) as s:
        s.set_psi(psi0)
        s.poisson()
        s.run(sched)
        fresh = s.get_psi()
        # stale potential from another field on entry
        s.set_psi(make_psi(n, 98, cfg.mfdm_solar))
        s.poisson()
        s.set_psi(psi0)
        s.run_checkpointed(sched, 2)
        psiT = s.get_psi()
        assert rel_l2(psiT, fresh) < 1e-14
        seed = _seed(psiT, rho_t, rbar)
        other = [(a * 1.01, b, c) for a, b, c in sched]
        with pytest.raises(F.IfdmError, match="schedule differs"):
            s.adjoint(other, seed)
        g = s.adjoint(sched, seed)
        assert np.isfinite(g).all()
        with pytest.raises(F.IfdmError, match="one reverse sweep"):
            s.adjoint(sched, seed)
        s.run_checkpointed(sched, 2)
        s.set_psi(psi0)
        with pytest.raises(F.IfdmError, match="one reverse sweep"):
            s.adjoint(sched, seed)


def test_adjoint_is_exact_for_a_pure_drift():
    """With dPE = dQI = 0 the step is the unitary drift: the adjoint of
    g_T is D^H g_T, i.e. the drift with -dKE applied to g_T."""
    import ifdm_code_amd as F
    n = 16
    cfg = F.Config.from_cli(2.5e-24, 0.0, 5.0, n)
    psi0 = make_psi(n, 2, cfg.mfdm_solar)
    rng = np.random.default_rng(1)
    gT = rng.standard_normal(psi0.shape) + 1j * rng.standard_normal(psi0.shape)
    sched = [(3e4, 0.0, 0.0), (2e4, 0.0, 0.0)]
    with F.Solver(cfg) as s:
        s.set_psi(psi0)
        s.poisson()
        s.run_checkpointed(sched, 1)
        g0 = s.adjoint(sched, gT)
        s.set_psi(gT)
        s.drift(-2e4)
        s.drift(-3e4)
        expect = s.get_psi()
    assert rel_l2(g0, expect) < 1e-13


TORCH_SCRIPT = r"""
import sys
sys.path.insert(0, sys.argv[1])
import torch                      # torch's HIP runtime first: the library then shares it
torch.cuda.init()
import numpy as np
import ifdm_code_amd as F
from ifdm_code_amd.autograd import evolve
from tests.test_gpu_adjoint import _setup, _seed
from tests.fields import rel_l2
n, steps = 16, 6
F_, O, rp, cfg, psi0, sched, rho_t, rbar = _setup(n, steps, 0.0, seed=21)
from tests.fields import make_psi
with F.Solver(cfg) as s:
    # the solver holds another field and its potential: evolve() must not
    # depend on them (it solves Poisson for psi0 itself)
    s.set_psi(make_psi(n, 99, cfg.mfdm_solar))
    s.poisson()
    p0 = torch.tensor(psi0, dtype=torch.complex128, device="cuda:0", requires_grad=True)
    pT = evolve(p0, s, sched, 3)
    rt = torch.tensor(rho_t, device="cuda:0")
    loss = (((pT.abs() ** 2) - rt) / rbar).pow(2).sum()
    loss.backward(retain_graph=True)
    g_torch = p0.grad.detach().cpu().numpy().copy()
    # a second backward through the same graph has no checkpoints left: it
    # must raise rather than return another function's gradient
    try:
        loss.backward()
        second = "returned"
    except RuntimeError as e:
        second = "raised"
    s.set_psi(psi0)
    s.run_checkpointed(sched, 3)
    psiT = s.get_psi()
    g_lib = s.adjoint(sched, _seed(psiT, rho_t, rbar))
# the same forward on a fresh solver
with F.Solver(cfg) as s2:
    s2.set_psi(psi0)
    s2.poisson()
    s2.run(sched)
    fresh = s2.get_psi()
e0 = rel_l2(pT.detach().cpu().numpy(), fresh)
e1 = rel_l2(pT.detach().cpu().numpy(), psiT)
e2 = rel_l2(g_torch, g_lib)
print("forward vs fresh", e0, "forward", e1, "gradient", e2, "second backward", second, flush=True)
assert e0 < 1e-14 and e1 < 1e-14 and e2 < 1e-12
assert second == "raised"
"""


def test_torch_autograd_function(tmp_path):
    """torch.autograd through the library (run in a fresh process that
    imports torch before the library: both use a HIP runtime with soname
    libamdhip64.so.7, and the first one loaded serves the process)."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = tmp_path / "torch_check.py"
    script.write_text(TORCH_SCRIPT)
    r = subprocess.run([sys.executable, str(script), root], capture_output=True, text=True, timeout=600, cwd=root)
    print(r.stdout[-2000:], r.stderr[-3000:])
    assert r.returncode == 0


def test_mixed_adjoint_against_fp64():
    """The mixed-precision adjoint (FP32 fields, power-of-two scaled psi and
    g, FP64 phases) against the FP64 adjoint on the same schedule, for a
    cold field (density rms 1%, phase rms 0.05 rad).

    * A linear loss L = Re sum conj(c) psi_T (seed c, fixed): both sweeps
      get the same seed and differ by the FP32 round-off of the mixed
      trajectory: its potential carries an error of about 1e-4 relative to
      the potential's fluctuation (FP32 density relative to a 1% contrast),
      which enters every reverse kick phase and accumulates over the 10
      steps (bound 5e-3).
    * The density loss with each path's own seed: the seed
      4(rho_T - rho_tgt) psi_T differences two nearly equal densities, so the
      mixed forward's error enters it amplified by about
      rho / |rho_T - rho_tgt|; the gradient may not amplify that seed error
      further (factor <= 10)."""
    import ifdm_code_amd as F
    from oracle import oracle as O
    n, steps = 32, 10
    rp = O.RunParams(n=n)
    cfg64 = F.Config.from_cli(2.5e-24, 0.0, 5.0, n)
    psi0 = make_psi(n, 8, cfg64.mfdm_solar, delta_rms=0.01, phase_rms=0.05)
    _st, used = O.simulate(rp, psi0, 1.0 / 51, steps)
    sched = [u[:3] for u in used]
    tgt, _ = O.simulate(rp, make_psi(n, 18, cfg64.mfdm_solar, delta_rms=0.01, phase_rms=0.05), 1.0 / 51, steps)
    rho_t = np.abs(tgt.psi) ** 2
    rbar = float(np.mean(np.abs(psi0) ** 2))
    rng = np.random.default_rng(11)
    c_lin = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))

    def grad(prec, seed_fn):
        c = F.Config.from_cli(2.5e-24, 0.0, 5.0, n, precision=prec)
        with F.Solver(c) as s:
            s.set_psi(psi0)
            s.poisson()
            s.run_checkpointed(sched, 5)
            seed = seed_fn(s.get_psi())
            return s.adjoint(sched, seed), seed

    g64_lin, _ = grad(0, lambda p: c_lin)
    gmx_lin, _ = grad(1, lambda p: c_lin)
    g64, seed64 = grad(0, lambda p: _seed(p, rho_t, rbar))
    gmx, seed_mx = grad(1, lambda p: _seed(p, rho_t, rbar))
    err_lin = rel_l2(gmx_lin, g64_lin)
    err_seed = rel_l2(seed_mx, seed64)
    err_full = rel_l2(gmx, g64)
    _say(f"   mixed adjoint: linear loss {err_lin:.3e}; density loss seed error {err_seed:.3e}, "
         f"gradient error {err_full:.3e}")
    assert err_lin < 5e-3
    assert err_full < 10 * err_seed + 1e-4


def test_adjoint_zeldovich_128_against_finite_differences():
    """128^3 FP64 from the Zel'dovich particle IC (box 10 Mpc/h, z = 50,
    seed 1234), 200 steps (the adaptive steps to a > 0.5, each split in 3),
    checkpoints every 20 steps; L = sum((rho_T - rho_tgt)/rho_bar)^2 with
    rho_tgt the final density from the seed-4321 IC on the same schedule.
    The gradient matches central finite differences (eps = 1e-6) along 16
    seeded directions to 1e-6."""
    import ifdm_code_amd as F
    from ifdm_code_amd import icgen
    n, total, K = 128, 200, 20
    _say(f"[adjoint zeldovich] n={n} steps={total} K={K}")
    cfg = F.Config.from_cli(2.5e-24, 0.0, 5.0, n)

    def ic(seed):
        pos, vel, pm, a0 = icgen.zeldovich(n, 10.0, z=50.0, seed=seed)
        with F.Solver(cfg) as s:
            s.ic_from_particles([(pos, vel)], 10.0e3, pm * 1e10, a0)
            return s.get_psi()

    psi0 = ic(1234)
    ad, a = [], 1.0 / 51
    with F.Solver(cfg) as s:
        s.set_psi(psi0)
        while a <= 0.5 and len(ad) < total:
            a, d = s.simulate_native(a, 1)
            ad += d
    m = -(-total // len(ad))
    sched = [tuple(x / m for x in d) for d in ad for _ in range(m)][:total]
    with F.Solver(cfg) as s:
        s.set_psi(ic(4321))
        s.poisson()
        s.run(sched)
        rho_t = np.abs(s.get_psi()) ** 2
    rbar = float(np.mean(np.abs(psi0) ** 2))
    with F.Solver(cfg) as s:
        def fwd(p):
            s.set_psi(p)
            s.poisson()
            s.run(sched)
            return s.get_psi()
        s.set_psi(psi0)
        s.poisson()
        s.run_checkpointed(sched, K)
        g0 = s.adjoint(sched, _seed(s.get_psi(), rho_t, rbar))
        rng = np.random.default_rng(2024)
        worst = 0.0
        for k in range(16):
            d = (rng.standard_normal(psi0.shape) + 1j * rng.standard_normal(psi0.shape)) * np.sqrt(rbar / 2)
            fd = (_loss(fwd(psi0 + 1e-6 * d), rho_t, rbar) - _loss(fwd(psi0 - 1e-6 * d), rho_t, rbar)) / 2e-6
            adj = float(np.sum((np.conj(g0) * d).real))
            worst = max(worst, abs(adj - fd) / abs(fd))
            _say(f"   dir {k}: adjoint {adj:.12e} fd {fd:.12e}")
        _say(f"   worst relative difference {worst:.3e}")
        assert worst <= 1e-6
