"""Benchmark: Schrodinger-Poisson KDK steps per second on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512] [--no-cpu-baseline]

One step is the reference's time-loop body (src/CosmoBEC_MPI_stan.cpp:
212-304 without snapshot I/O): Poisson reductions -> Timestep_criteria (FP64
host quadrature) -> Split_op_MPI (kick, 3-D FFT drift, Poisson, kick), run by
libifdm.so's ifdm_simulate with adjacent half-kicks merged.  Inputs are
resident in HBM when the timed region starts.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "KDK steps/sec at 512^3: FP64 vs FP32-mixed at stated error; % HBM roofline"
HBM_PEAK_GBS = 8000.0
PASS_NAMES = ["T_drift_table", "A_kick_ydrift", "B_xdrift", "C_zdrift_src_r2c",
              "F_yfft_half", "G_xfft_green_xifft_half", "H_yifft_half", "I_c2r_z", "reduce"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pass_bytes(n: int, sc: int = 16, sr: int = 8) -> list[float]:
    """Algorithmic HBM bytes of each pass: one read and one write of each
    field it touches; s_c complex, s_r real element size; h = (N/2+1)/N.
    A, B, C each run one axis's drift round trip on psi (the drift factor
    separates by axis); A also reads phi for the kick, C also writes the
    source's half spectrum."""
    n3 = float(n) ** 3
    h = (n // 2 + 1) / n
    return [0.0, n3 * (2 * sc + sr), n3 * 2 * sc, n3 * (2 * sc + h * sc),
            n3 * 2 * h * sc, n3 * 2 * h * sc, n3 * 2 * h * sc, n3 * (h * sc + sr), 0.0]


def pencil_model_bytes(n: int, sc: int = 16) -> float:
    """B_alg of the minimal five-pass-drift pencil schedule, (11 + 8h) s_c N^3:
    the reference point of the step's HBM roofline fraction."""
    h = (n // 2 + 1) / n
    return (11 + 8 * h) * sc * float(n) ** 3


def make_ic(F, n: int, xmax_in: float, seed: int = 1234) -> np.ndarray:
    """The configs' IC: n^3 particles on a lattice with Zel'dovich
    displacements (P(k) ~ k^-2.5 to Nyquist, rms 0.2 cells, z = 50), CIC
    deposited by the library into psi = 1.225 sqrt(rho/m_s) e^{i theta}
    (the reference's IC path).  Returns psi (FP64, host)."""
    from ifdm_code_amd import icgen
    box = 2 * xmax_in * 1e3                                  # kpc/h
    pos, vel, pm, a = icgen.zeldovich(n, 2 * xmax_in, z=50.0, seed=seed, progress=lambda m: log(f"[bench] {m}"))
    cfg = F.Config.from_cli(2.5e-24, 0.0, xmax_in, n)
    with F.Solver(cfg) as s:
        s.ic_from_particles([(pos, vel)], box, pm * 1e10, a)
        del pos, vel
        return s.get_psi()


def cpu_baseline(n: int, psi: np.ndarray, cfg_g_in: float, xmax_in: float, steps: int = 5, warm: int = 2,
                 library_fft: bool = False) -> dict:
    """The reference step on the host's cores, restated by the oracle:
    `warm` untimed steps, then the median of `steps` timed ones.  Default:
    the C restatement (plain radix-2 FFT per line, OpenMP); library_fft:
    the same step over scipy.fft (pocketfft) with one worker per host
    thread (oracle.FastState)."""
    from oracle import oracle as O
    rp = O.RunParams(n=n, mfdm=2.5e-24, g_in=cfg_g_in, xmax_in=xmax_in)
    st = O.FastState(rp, psi) if library_fft else O.State(rp, psi)
    a = 1.0 / 51
    times = []
    for i in range(warm + steps):
        t0 = time.perf_counter()
        st.poisson()
        m = st.maxima()
        dKE, dPE, dQI, a, _ = O.timestep_criteria(rp, a, m["max_phi"], m["max_rho"])
        st.split_op(dKE, dPE, dQI)
        times.append(time.perf_counter() - t0)
        log(f"[cpu baseline] step {i} {times[-1]:.2f} s")
    t = float(np.median(times[warm:]))
    if library_fft:
        cores = st.workers
        st.close()
        how = ("the oracle's restatement over scipy.fft (pocketfft, SIMD mixed-radix transforms, one worker per "
               "host thread; elementwise work split over x-planes in a thread pool)")
    else:
        cores = O.num_threads()
        how = ("the oracle's C restatement, each 3-D FFT a plain radix-2 FFT per line with OpenMP over lines "
               "(not an FFTW-class transform)")
    return {"value": 1.0 / t, "unit": "KDK steps/s", "cores": cores, "kind": "port",
            "sample": f"median of {steps} timed steps after {warm} warm-up steps of the same {n}^3 FP64 workload "
                      f"(poisson + Timestep_criteria + Split_op_MPI, 6 3-D FFTs per step) with {how}",
            "step_s": [round(x, 3) for x in times[warm:]]}


def small_mesh(F, steps: int = 200, cpu_steps: int = 10) -> dict:
    """The 64^3 FP64 workload (Zel'dovich IC, box 10 Mpc/h, z = 50): the
    library's adaptive loop on the GPU (`steps` steps resuming a run after 5
    warm-up steps, launch-latency bound at this size) and the oracle's C
    restatement on the host cores (median of `cpu_steps` steps)."""
    n = 64
    psi = make_ic(F, n, 5.0)
    with F.Solver(F.Config.from_cli(2.5e-24, 0.0, 5.0, n)) as s:
        s.set_psi(psi)
        a, _ = s.simulate_native(1.0 / 51, 5, keep_pending=True)
        t0 = time.perf_counter()
        s.simulate_native(a, steps, resume=True, keep_pending=True)
        gpu = steps / (time.perf_counter() - t0)
        s.flush()
    cpu = cpu_baseline(n, psi, 0.0, 5.0, cpu_steps)
    return {"workload": "KDK step at 64^3, box 10 Mpc/h, z=50, mfdm=2.5e-24 eV, g=0, adaptive dt, FP64",
            "gpu_steps_per_s": round(gpu, 2), "gpu_steps": steps, "cpu_baseline": cpu}


def rel_l2(a, b) -> float:
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def timed_adaptive(F, cfg, psi, warmup: int, steps: int, prof_steps: int, keep_state: bool, persistent: int = -1):
    """`warmup` untimed adaptive steps (ifdm_simulate_ex), the last of them (up
    to 3) with HIP events around every pass to find the dominant pass; then
    `steps` timed steps carrying events around that pass only (its live
    duration over the timed region); then `prof_steps` more with events
    around every pass for the per-pass table.  Returns (steps/s, seconds, avg
    ms per pass from the last phase, schedule of the warm-up + timed steps,
    psi and phi after them (or None), a, dominant pass index, its avg ms over
    the timed steps, the host copy times of psi: set_psi (host -> HBM, the
    second of two calls) and get_psi (HBM -> host, after the run))."""
    s = F.Solver(cfg)
    if persistent >= 0:
        s.set_option("persistent", persistent)
    s.set_psi(psi)
    t0 = time.perf_counter()
    s.set_psi(psi)
    t_set = time.perf_counter() - t0
    # one run in pieces (ifdm_simulate_ex): the entry Poisson solve belongs
    # to the first warm-up piece, each later piece resumes the run, so the
    # timed piece holds exactly `steps` steps of the run's steady state
    nsel = min(3, warmup)
    started = warmup - nsel > 0
    a, sched_w = s.simulate_native(1.0 / 51, warmup - nsel, keep_pending=True) if started else (1.0 / 51, [])
    dom = 1                                   # pass A unless the warm-up profile says otherwise
    if nsel:
        s.set_profiling(True)
        a, d = s.simulate_native(a, nsel, resume=started, keep_pending=True)
        started = True
        sched_w += d
        ms_sum, cnt = s.get_profile()
        dom = int(np.argmax(ms_sum[:-1]))
    s.set_option("profile_mask", 1 << dom)
    s.set_profiling(True)
    t0 = time.perf_counter()
    a, sched_t = s.simulate_native(a, steps, resume=started, keep_pending=True)
    el = time.perf_counter() - t0
    ms_sum, cnt = s.get_profile()
    dom_ms = ms_sum[dom] / max(cnt, 1)
    psi_out = s.get_psi() if keep_state else None      # (applies the pending half-kick)
    phi_out = s.get_phi() if keep_state else None
    s.set_option("profile_mask", (1 << F.NPASS) - 1)
    s.set_profiling(True)
    s.simulate_native(a, prof_steps)
    ms_sum, cnt = s.get_profile()
    s.set_profiling(False)
    t0 = time.perf_counter()
    s.get_psi()
    t_get = time.perf_counter() - t0
    s.close()
    avg = [m / max(cnt, 1) for m in ms_sum]
    return steps / el, el, avg, sched_w + sched_t, psi_out, phi_out, a, dom, dom_ms, (t_set, t_get)


def host_io(t, step_s, n):
    """psi's host copies (complex128 host arrays in the reference layout, the
    ABI's set_psi / get_psi) beside the step time: the rate of a run of k
    steps that starts and ends with psi in host memory.  Not the headline
    value, which keeps psi resident in HBM."""
    t_set, t_get = t
    nb = 16 * n ** 3
    return {"bytes_each_way": nb, "set_psi_s": round(t_set, 4), "get_psi_s": round(t_get, 4),
            "set_psi_GBs": round(nb / t_set / 1e9, 2), "get_psi_GBs": round(nb / t_get / 1e9, 2),
            "steps_per_s_incl_host_copies": {str(k): round(k / (t_set + k * step_s + t_get), 3)
                                             for k in (1, 100, 1000)}}


def _phase_errors(psi, psi64, w):
    """(rms, global): the density-weighted rms of the wrapped phase
    difference after removing its weighted circular mean (a global phase,
    which no observable depends on), and that mean."""
    d = psi * np.conj(psi64)
    z = np.sum(w * d / np.maximum(np.abs(d), 1e-300))
    g = float(np.angle(z))
    dp = np.angle(d * complex(np.cos(-g), np.sin(-g)))
    return float(np.sqrt(np.sum(w * dp * dp))), g


def _field_errors(s, psi64, phi64, pk64, norm64) -> dict:
    """Per-field errors of solver s's state against the FP64 run's."""
    psi = s.get_psi()
    st = s.poisson()
    phi = s.get_phi()
    _k, pk, _nm = s.power_spectrum()
    rho64, rho = np.abs(psi64) ** 2, np.abs(psi) ** 2
    w = rho64 / rho64.sum()
    ph, gph = _phase_errors(psi, psi64, w)
    good = pk64 > 0
    rel = np.abs(pk[good] / pk64[good] - 1) if good.any() else np.array([np.nan])
    return {"density_rel_l2": rel_l2(rho, rho64),
            "phase_rho_weighted_rms_rad": ph, "global_phase_rad": gph,
            "potential_rel_l2": rel_l2(phi, phi64),
            "pk_max_rel": float(rel.max()), "pk_median_rel": float(np.median(rel)),
            "norm_rel": abs(st["norm"] / norm64 - 1)}


def round_c64(psi: np.ndarray) -> np.ndarray:
    """psi rounded to complex64 precision (through an exact power-of-two
    scaling, since |psi| exceeds the FP32 range)."""
    e = int(np.frexp(np.max(np.abs(psi)))[1])
    f = 2.0 ** -e
    return (psi * f).astype(np.complex64).astype(np.complex128) / f


def split_schedule(F, cfg, psi, total: int, a_stop: float = 0.5, mode: str = "split"):
    """The FP64 adaptive steps (Timestep_criteria each step) up to the
    reference's own end of run, the first step that takes a past a_stop
    (src/CosmoBEC_MPI_stan.cpp:250-255).  When `total` asks for more steps
    than that: mode "split" divides each into m equal parts (D_KE, D_PE,
    D_QI divided by m) so the same expansion span is covered; mode "hold"
    repeats the last step's integrals until `total` (the run then continues
    past a_stop).  Returns (the first `total` steps, adaptive step count, a at
    the end of the adaptive run, m; m = 1 for "hold")."""
    sched, a = [], 1.0 / 51
    with F.Solver(cfg) as sg:
        sg.set_psi(psi)
        while a <= a_stop and len(sched) < total:
            a, d = sg.simulate_native(a, 1)
            sched += d
    if mode == "hold":
        return (sched + [sched[-1]] * (total - len(sched)))[:total], len(sched), a, 1
    m = max(1, -(-total // len(sched)))
    return [tuple(x / m for x in d) for d in sched for _ in range(m)][:total], len(sched), a, m


def error_study(F, cfg, cfgm, psi, checkpoints, a_stop: float = 0.5, mode: str = "split") -> dict:
    """Precision error against the FP64 GPU run of the same IC.

    Schedule: the FP64 adaptive run (Timestep_criteria each step) up to the
    reference's own end of run, the first step that takes a past 0.5
    (src/CosmoBEC_MPI_stan.cpp:250-255).  When the largest checkpoint asks
    for more steps than that, every step is split into m equal parts
    (D_KE, D_PE, D_QI each divided by m) so the same expansion span is
    covered in that many steps (mode "split"; mode "hold" repeats the last
    adaptive step instead, past a = 0.5).  The FP64 path, the mixed path and an FP64
    run started from the IC rounded to complex64 (the error the rounding of
    the IC alone grows into: the floor the dynamics set) replay that
    schedule; at each checkpoint (cumulative step counts) the per-field
    errors of the latter two against the FP64 path are recorded."""
    steps, nad, a, m = split_schedule(F, cfg, psi, max(checkpoints), a_stop, mode)
    cps = sorted({min(c, len(steps)) for c in checkpoints})
    log(f"[bench] error study: {nad} adaptive steps to a = {a:.4f}, {mode} (m = {m}) -> {len(steps)} steps")
    s64, smx, sfl = F.Solver(cfg), F.Solver(cfgm), F.Solver(cfg)
    try:
        for s, p in ((s64, psi), (smx, psi), (sfl, round_c64(psi))):
            s.set_psi(p)
            s.poisson()
        done, rows = 0, []
        for c in cps:
            for s in (s64, smx, sfl):
                s.run(steps[done:c])
            done = c
            psi64 = s64.get_psi()
            norm64 = s64.poisson()["norm"]
            phi64 = s64.get_phi()
            _k, pk64, _nm = s64.power_spectrum()
            row = {"steps": c, "mixed": _field_errors(smx, psi64, phi64, pk64, norm64),
                   "fp64_from_c64_rounded_ic": _field_errors(sfl, psi64, phi64, pk64, norm64)}
            log(f"[bench] error after {c} steps: {json.dumps(row)}")
            rows.append(row)
            del psi64, phi64
        return {"adaptive_steps": nad, "a_end": a, "schedule": mode, "split": m, "steps": len(steps),
                "checkpoints": rows}
    finally:
        for s in (s64, smx, sfl):
            s.close()


def _hip_free_bytes() -> int:
    """Free device memory (hipMemGetInfo) of the current device."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        raise RuntimeError("hipMemGetInfo failed")
    return free.value


def adjoint_study(F, n: int, xmax_in: float, g_in: float, steps: int, K: int, fd_dirs: int) -> dict:
    """Gradient of L = sum((rho_T - rho_tgt) / rho_bar)^2 with respect to
    psi_0 (rho_tgt: the final density from the Zel'dovich IC of seed 4321
    on the same schedule) by the FP64 and the mixed adjoint with device
    checkpoints every K steps: time, device memory held by the context
    after the sweep (every buffer it owns), the mixed gradient's error
    against the FP64 one, and the FP64 gradient against central finite
    differences (eps = 1e-6) along fd_dirs seeded directions."""
    psi0 = make_ic(F, n, xmax_in)
    cfg = F.Config.from_cli(2.5e-24, g_in, xmax_in, n)
    sched, nad, a_end, m = split_schedule(F, cfg, psi0, steps)
    log(f"[bench] adjoint study: {len(sched)} steps ({nad} adaptive to a = {a_end:.4f}, split {m}x), K = {K}")
    with F.Solver(cfg) as s:
        s.set_psi(make_ic(F, n, xmax_in, seed=4321))
        s.poisson()
        s.run(sched)
        rho_t = np.abs(s.get_psi()) ** 2
    rbar = float(np.mean(np.abs(psi0) ** 2))

    def loss(p):
        return float(np.sum(((np.abs(p) ** 2 - rho_t) / rbar) ** 2))

    out, grads = {"steps": len(sched), "K": K, "adaptive_steps": nad, "split": m}, {}
    for name, prec in (("fp64", F.FP64), ("mixed", F.MIXED)):
        free0 = _hip_free_bytes()
        c = F.Config.from_cli(2.5e-24, g_in, xmax_in, n, precision=prec)
        with F.Solver(c) as s:
            s.set_psi(psi0)
            s.poisson()
            t0 = time.perf_counter()
            s.run_checkpointed(sched, K)
            tf = time.perf_counter() - t0
            psiT = s.get_psi()
            seed = 4 * (np.abs(psiT) ** 2 - rho_t) / rbar ** 2 * psiT
            L = loss(psiT)
            del psiT
            t0 = time.perf_counter()
            grads[name] = s.adjoint(sched, seed)
            ta = time.perf_counter() - t0
            held = free0 - _hip_free_bytes()
            del seed
            r = {"forward_s": round(tf, 3), "adjoint_s": round(ta, 3), "device_bytes_held": held, "loss": L}
            if name == "fp64" and fd_dirs:
                rng = np.random.default_rng(77)
                rels = []
                for _ in range(fd_dirs):
                    d = (rng.standard_normal(psi0.shape) + 1j * rng.standard_normal(psi0.shape)) * np.sqrt(rbar / 2)
                    lp = []
                    for sgn in (1, -1):
                        s.set_psi(psi0 + sgn * 1e-6 * d)
                        s.poisson()
                        s.run(sched)
                        lp.append(loss(s.get_psi()))
                    fd = (lp[0] - lp[1]) / 2e-6
                    ad = float(np.sum((np.conj(grads[name]) * d).real))
                    rels.append(abs(ad - fd) / abs(fd))
                    log(f"[bench] adjoint vs FD: {ad:.12e} {fd:.12e}")
                    del d
                r["fd_rel_err_max"] = max(rels)
            log(f"[bench] adjoint {name}: {json.dumps(r)}")
            out[name] = r
    out["mixed_grad_rel_l2_vs_fp64"] = rel_l2(grads["mixed"], grads["fp64"])
    return out


def counter_traffic(n: int, precision: str) -> dict:
    """HBM bytes per dispatch of each pass from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (profiles/kdk_<n>_<precision>_hbm_traffic.md),
    already scaled by the copy-kernel calibration; {} when absent."""
    path = os.path.join(ROOT, "profiles", f"kdk_{n}_{precision}_hbm_traffic.md")
    out = {}
    if os.path.exists(path):
        txt = open(path).read()
        if "```traffic" in txt:
            for line in txt.split("```traffic", 1)[1].split("```", 1)[0].strip().splitlines():
                name, prec, val = line.split()
                out[name] = float(val)
    return out


def roofline(n: int, dom: int, dom_ms: float, avg: list, sc: int, sr: int, precision: str) -> dict:
    """The dominant pass: its algorithmic bytes over its average duration
    from HIP events around it during the timed steps (dom_ms); beside it the
    average from the steps profiled pass by pass afterwards."""
    pb = pass_bytes(n, sc, sr)
    achieved = pb[dom] / (dom_ms * 1e-3) / 1e9
    traffic = counter_traffic(n, precision).get(PASS_NAMES[dom])
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": PASS_NAMES[dom],
            "kernel_avg_ms": round(dom_ms, 4), "kernel_avg_ms_every_pass_profiled": round(avg[dom], 4),
            "kernel_alg_bytes": pb[dom]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=160)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--prof-steps", type=int, default=10)
    ap.add_argument("--no-small", action="store_true", help="skip the 64^3 GPU and CPU timing")
    ap.add_argument("--error-steps", type=str, default="100,1000",
                    help="comma list of cumulative step counts at which the mixed path's error against the "
                         "FP64 run is reported ('' to skip)")
    ap.add_argument("--error-schedule", choices=("split", "hold"), default="split",
                    help="past the adaptive run's end at a = 0.5: split every step (default) or hold the last one")
    ap.add_argument("--no-mixed-check", action="store_true",
                    help="skip the mixed path's replay of the timed FP64 schedule (host memory at 1024^3)")
    ap.add_argument("--adjoint-steps", type=int, default=0,
                    help="also run the adjoint study over this many steps (0: skip)")
    ap.add_argument("--adjoint-k", type=int, default=50, help="checkpoint interval of the adjoint study")
    ap.add_argument("--adjoint-fd", type=int, default=2, help="finite-difference directions of the adjoint study")
    ap.add_argument("--persistent", type=int, default=-1,
                    help="persistent strided-pass mask for both precisions (-1: the library's defaults)")
    ap.add_argument("--plane-pad", type=int, default=0,
                    help="device cells after each x-plane of psi and phi (0: the library's choice, <0: none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU: this build has no multi-GPU path (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import ifdm_code_amd as F
    n = args.n
    xmax_in = {512: 10.0, 1024: 20.0}.get(n, 5.0)   # box 20 / 40 Mpc/h at 512^3 / 1024^3, else 10 Mpc/h
    cfg0 = F.Config.from_cli(2.5e-24, 0.0, xmax_in, n)
    log(f"[bench] building synthetic IC n={n}")
    t0 = time.perf_counter()
    psi = make_ic(F, n, xmax_in)
    log(f"[bench] IC ready in {time.perf_counter() - t0:.1f} s")

    # g sized so that g <|psi|^2> = 1% of m rms(phi_IC) (self-interaction vs gravity)
    with F.Solver(cfg0) as s0:
        s0.set_psi(psi)
        st = s0.poisson()
        phi = s0.get_phi()
        rms_phi = float(np.sqrt(np.mean(phi * phi)))
        mean_rho = st["norm"] / n ** 3
        del phi
        copy_gbs = s0.stream_copy(2 << 30, 10)
    log(f"[bench] stream copy {copy_gbs:.0f} GB/s")
    mt = cfg0.mfdm / F.constant("c") ** 2
    g = 0.01 * mt * rms_phi / mean_rho              # eV Mpc^3
    g_in = g * F.constant("cm2Mpc") ** 3            # eV cm^3, the CLI unit
    cfg = F.Config.from_cli(2.5e-24, g_in, xmax_in, n, plane_pad=args.plane_pad)
    log(f"[bench] g = {g_in:.4e} eV cm^3")

    sps, el, avg, sched, psi64, phi64, a, dom, dom_ms, hio = timed_adaptive(F, cfg, psi, args.warmup, args.steps,
                                                                       args.prof_steps,
                                                                       not (args.no_mixed or args.no_mixed_check),
                                                                       args.persistent)
    log(f"[bench] FP64: {args.steps} steps in {el:.3f} s ({sps:.2f} steps/s), a = {a:.5f}")
    pb = pass_bytes(n)
    b_alg = sum(pb)
    step_gbs = b_alg * sps / 1e9
    out = {
        "metric": METRIC, "value": round(sps, 3), "unit": "KDK steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * el / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: Zel'dovich particle IC (lattice, P(k)~k^-2.5 to Nyquist, rms displacement 0.2 cells, "
                "z=50, seed 1234) CIC-deposited into psi by the library",
        "config": {"workload": f"KDK step at {n}^3, box {2 * xmax_in:g} Mpc/h, z=50, mfdm=2.5e-24 eV, "
                               f"g={g_in:.3e} eV cm^3 (1% of gravity), adaptive dt, FP64",
                   "n": n, "precision": "fp64"},
        "roofline": roofline(n, dom, dom_ms, avg, 16, 8, "fp64"),
        # frac_of_8TBs: the step on the reference five-pass-drift schedule's
        # bytes, (11 + 8h) s_c N^3 per step at 8 TB/s; bytes_moved_*: this
        # schedule's own algorithmic bytes (fewer: three drift passes)
        "step_hbm": {"alg_bytes_per_step": pencil_model_bytes(n), "frac_of_8TBs":
                         round(pencil_model_bytes(n) * sps / 1e9 / HBM_PEAK_GBS, 4),
                     "bytes_moved_per_step": b_alg, "achieved_GBs": round(step_gbs, 1),
                     "bytes_moved_frac_of_8TBs": round(step_gbs / HBM_PEAK_GBS, 4),
                     "cell_updates_per_s": round(sps * n ** 3, 1),
                     "stream_copy_GBs": round(copy_gbs, 1)},
        "passes_ms": {k: round(v, 4) for k, v in zip(PASS_NAMES, avg)},
        "host_io": host_io(hio, el / args.steps, n),
    }

    if not args.no_mixed:
        cfgm = F.Config.from_cli(2.5e-24, g_in, xmax_in, n, precision=F.MIXED, plane_pad=args.plane_pad)
        spsm, elm, avgm, _schm, _p, _f, am, domm, domm_ms, hiom = timed_adaptive(F, cfgm, psi, args.warmup, args.steps,
                                                                          args.prof_steps, False, args.persistent)
        log(f"[bench] mixed: {args.steps} steps in {elm:.3f} s ({spsm:.2f} steps/s)")
        pbm = pass_bytes(n, 8, 4)
        out["mixed"] = {
            "value": round(spsm, 3), "unit": "KDK steps/s", "ms_per_step": round(1e3 * elm / args.steps, 4),
            "speedup_vs_fp64": round(spsm / sps, 3), "dtype": "c64 storage/FP32 FFT, FP64 phases+sums",
            "roofline": roofline(n, domm, domm_ms, avgm, 8, 4, "mixed"),
            "step_hbm_frac_of_8TBs": round(pencil_model_bytes(n, 8) * spsm / 1e9 / HBM_PEAK_GBS, 4),
            "bytes_moved_frac_of_8TBs": round(sum(pbm) * spsm / 1e9 / HBM_PEAK_GBS, 4),
            "passes_ms": {k: round(v, 4) for k, v in zip(PASS_NAMES, avgm)},
            "host_io": host_io(hiom, elm / args.steps, n),
        }
    if not args.no_mixed and not args.no_mixed_check:
        # error: the mixed path replays the FP64 run's own adaptive schedule
        # (warm-up + timed steps, unsplit) from the same IC; beside it an FP64
        # replay from the IC rounded to complex64, the error that rounding
        # alone grows into over the same steps
        def replay(cfg_r, psi_r):
            with F.Solver(cfg_r) as sr:
                sr.set_psi(psi_r)
                sr.poisson()
                sr.run([tuple(x) for x in sched])
                p = sr.get_psi()
                sr.poisson()
                return p, sr.get_phi()

        rho64 = np.abs(psi64) ** 2
        w = rho64 / rho64.sum()
        res = {}
        for key, (cfg_r, psi_r) in (("mixed", (cfgm, psi)), ("fp64_from_c64_rounded_ic", (cfg, round_c64(psi)))):
            p, f = replay(cfg_r, psi_r)
            ph, _g = _phase_errors(p, psi64, w)
            res[key] = {"density_rel_l2": rel_l2(np.abs(p) ** 2, rho64), "phase_rho_weighted_rms_rad": ph,
                        "potential_rel_l2": rel_l2(f, phi64)}
            del p, f
        out["mixed"]["error_vs_fp64_after_steps"] = {
            "steps": len(sched), "a_end": round(a, 5),
            "schedule": "the FP64 run's own adaptive steps (warm-up + timed), unsplit",
            **res["mixed"], "fp64_from_c64_rounded_ic": res["fp64_from_c64_rounded_ic"]}
        del psi64, rho64, w, phi64
    cps = [int(x) for x in args.error_steps.split(",") if x.strip()]
    if not args.no_mixed and cps:
        out["mixed"]["error_vs_fp64"] = error_study(F, cfg, cfgm, psi, sorted(cps), mode=args.error_schedule)
    if not args.no_mixed:
        # the mixed path's density error in one place: on the FP64 run's own
        # adaptive steps (unsplit: the reference's step sizes) and after the
        # 1000-step split schedule
        m = out["mixed"]
        head = {}
        if "error_vs_fp64_after_steps" in m:
            head["unsplit_steps"] = m["error_vs_fp64_after_steps"]["steps"]
            head["density_rel_l2_unsplit"] = m["error_vs_fp64_after_steps"]["density_rel_l2"]
        for row in m.get("error_vs_fp64", {}).get("checkpoints", []):
            head[f"density_rel_l2_after_{row['steps']}_split_steps"] = row["mixed"]["density_rel_l2"]
        if head:
            m["density_error"] = head
    if args.adjoint_steps:
        out["adjoint"] = adjoint_study(F, n, xmax_in, g_in, args.adjoint_steps, args.adjoint_k, args.adjoint_fd)
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle over scipy.fft) ...")
        out["cpu_baseline"] = cpu_baseline(n, psi, g_in, xmax_in, args.cpu_steps, library_fft=True)
        log("[bench] CPU baseline (oracle C restatement, radix-2) ...")
        out["cpu_baseline_radix2"] = cpu_baseline(n, psi, g_in, xmax_in, args.cpu_steps)
        if not args.no_small:
            del psi
            log("[bench] 64^3: GPU adaptive loop and CPU oracle ...")
            out["mesh_64"] = small_mesh(F)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
