"""CPU oracle for the fuzzy-dark-matter KDK step -- TEST INFRASTRUCTURE ONLY.

This module restates the reference's FP64 algorithm (jchan192/IFDM_Code,
``src/CosmoBEC_MPI_stan.cpp``) on the CPU.  The array work runs in
``liboracle.so`` (plain C, ``oracle/ifdm_oracle.c``); the time-step
criteria and the scale-factor integrals run here, the integrals through
``scipy.integrate.quad`` -- QUADPACK's QAGS, the routine GSL's
``gsl_integration_qags`` re-implements.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module, and only as the checker / baseline.  The product
(``ifdm_code_amd``) never imports it.

Parity pins (each function -> the independent check in tests/):

=====================  ===================================================
oracle function        pinned by
=====================  ===================================================
fft3d_c2c / r2c / c2r  brute-force O(N^6) DFT at N<=8; numpy pocketfft at
                       N=16..32; Parseval; impulse -> constant
poisson                closed form: rho = rho0 (1 + eps cos k0.x) gives
                       phi = -4 pi G rho0 eps cos(k0.x) / |k0|^2
kick                   closed form: uniform psi with g != 0 (phase
                       g|psi|^2 dQI / hbar per two half-kicks); the
                       potential phase m phi dPE / (c^2 hbar) with the
                       analytic cosine phi; |psi| unchanged
drift                  closed form: plane-wave superposition evolves as
                       A_k exp(-i hbar k^2 dKE / (2 m))
split_op               norm conservation; time reversal (conjugate, step,
                       conjugate); Galilean plane wave
maxima                 brute-force numpy max / sum over the same entries
timestep_criteria      the kinetic / potential / self-interaction phase of
                       one step equals 2 pi in the da -> 0 limit (the
                       criterion's design); integrals vs closed-form
                       Einstein-de Sitter limit (Omega_m = 1)
ic_from_particles      mass conservation, uniform lattice -> uniform rho,
                       trilinear weights, periodic wrap, single velocity
                       mode -> analytic theta; the float grids (f32) hold
                       float values accumulated in particle order
Triggers               write sequences worked out by hand from :238-284
                       for a fixed a(tn) trajectory
=====================  ===================================================
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

# --- constants: src/Cosmoinput3D.h:2-17 (evaluated with the same libm pow)
m2kpc = 3.086e19
m2Mpc = 3.086e22
cm2Mpc = 3.086e24
s2Myr = 60.0 * 60.0 * 24.0 * 365.0 * 1e6
kg2eV = 1.783e-36
kg2solar = 1.989e30
rhoc_const = 3.51059e10
h_0 = 0.677
H_0 = h_0 * 100 * s2Myr / m2Mpc * 1e3
hbar = 6.58211928e-16 / s2Myr
Omega_m0 = 0.276
bigG = 6.674e-11 * kg2solar / math.pow(m2Mpc, 3) * math.pow(s2Myr, 2)
c_light = 2.99792458e8 / m2Mpc * s2Myr


class _Params(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("xmax", ctypes.c_double), ("mfdm", ctypes.c_double),
                ("mfdm_solar", ctypes.c_double), ("g", ctypes.c_double), ("rhoc", ctypes.c_double),
                ("hbar", ctypes.c_double), ("c", ctypes.c_double), ("bigG", ctypes.c_double)]


def lib():
    """Load liboracle.so (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        pp = ctypes.POINTER(_Params)
        L.oracle_fft3d_c2c.argtypes = [dp, ctypes.c_int, ctypes.c_int]
        L.oracle_fft3d_r2c.argtypes = [dp, dp, ctypes.c_int]
        L.oracle_fft3d_c2r.argtypes = [dp, dp, ctypes.c_int]
        L.oracle_poisson.argtypes = [pp, dp, dp, dp]
        L.oracle_kick.argtypes = [pp, dp, dp, ctypes.c_double, ctypes.c_double]
        L.oracle_drift.argtypes = [pp, dp, ctypes.c_double]
        L.oracle_split_op.argtypes = [pp, dp, dp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.oracle_maxima.argtypes = [pp, dp, dp, ctypes.c_int, dp]
        fp = ctypes.POINTER(ctypes.c_float)
        L.oracle_cic_chunk.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, fp, fp, ctypes.c_long,
                                       ctypes.c_int, ctypes.c_int, dp, dp, dp, dp]
        L.oracle_cic_chunk.restype = ctypes.c_long
        L.oracle_ic_assemble.argtypes = [pp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int, dp, dp, dp, dp, dp, dp]
        L.oracle_num_threads.restype = ctypes.c_int
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        _LIB = L
    return _LIB


def _dp(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class RunParams:
    """The run-time parameters of src/CosmoBEC_MPI_stan.cpp:329-339 after
    its unit conversions (mfdm in eV, g_in in eV cm^3, xmax_in = box/2 in
    Mpc/h)."""
    n: int
    mfdm: float = 2.5e-24
    g_in: float = 0.0
    xmax_in: float = 5.0
    rhoc: float = rhoc_const

    @property
    def mfdm_solar(self):            # :330
        return self.mfdm * kg2eV / kg2solar

    @property
    def g(self):                     # :331
        return self.g_in / math.pow(cm2Mpc, 3)

    @property
    def xmax(self):                  # :332
        return self.xmax_in / h_0

    @property
    def dx(self):                    # struct3D.h:52
        return 2.0 * self.xmax / self.n

    def c_params(self) -> _Params:
        return _Params(self.n, self.xmax, self.mfdm, self.mfdm_solar, self.g, self.rhoc,
                       hbar, c_light, bigG)


def kvec(n: int, xmax: float) -> np.ndarray:
    """src/struct3D.h:76-85"""
    i = np.arange(n, dtype=np.float64)
    return np.where(i < n // 2, i * math.pi / xmax, (i - n) * math.pi / xmax)


# ----------------------------------------------------------------- FFTs
def fft3d_c2c(a: np.ndarray, sign: int) -> np.ndarray:
    out = np.ascontiguousarray(a, dtype=np.complex128).copy()
    lib().oracle_fft3d_c2c(_dp(out.view(np.float64)), out.shape[0], int(sign))
    return out


def fft3d_r2c(real: np.ndarray) -> np.ndarray:
    n = real.shape[0]
    pad = np.zeros((n, n, 2 * (n // 2 + 1)))
    pad[:, :, :n] = real
    out = np.zeros((n, n, n // 2 + 1), np.complex128)
    lib().oracle_fft3d_r2c(_dp(pad), _dp(out.view(np.float64)), n)
    return out


def fft3d_c2r(half: np.ndarray) -> np.ndarray:
    n = half.shape[0]
    h = np.ascontiguousarray(half, dtype=np.complex128).copy()
    pad = np.zeros((n, n, 2 * (n // 2 + 1)))
    lib().oracle_fft3d_c2r(_dp(h.view(np.float64)), _dp(pad), n)
    return pad[:, :, :n].copy()


# ----------------------------------------------------------------- state
class State:
    """psi (N^3 complex128), phi in the reference's padded layout and the
    half-spectrum scratch, as the reference's ``Operators`` (struct3D.h:98)."""

    def __init__(self, rp: RunParams, psi: np.ndarray):
        n = rp.n
        self.rp = rp
        self.cp = rp.c_params()
        self.psi = np.ascontiguousarray(psi, dtype=np.complex128).reshape(n, n, n).copy()
        self.phi_pad = np.zeros((n, n, 2 * (n // 2 + 1)))
        self.phikk = np.zeros((n, n, n // 2 + 1), np.complex128)

    @property
    def phi(self) -> np.ndarray:
        return self.phi_pad[:, :, :self.rp.n].copy()

    def poisson(self):
        lib().oracle_poisson(ctypes.byref(self.cp), _dp(self.psi.view(np.float64)), _dp(self.phi_pad),
                             _dp(self.phikk.view(np.float64)))

    def kick(self, dPE: float, dQI: float):
        lib().oracle_kick(ctypes.byref(self.cp), _dp(self.psi.view(np.float64)), _dp(self.phi_pad), dPE, dQI)

    def drift(self, dKE: float):
        lib().oracle_drift(ctypes.byref(self.cp), _dp(self.psi.view(np.float64)), dKE)

    def split_op(self, dKE: float, dPE: float, dQI: float):
        lib().oracle_split_op(ctypes.byref(self.cp), _dp(self.psi.view(np.float64)), _dp(self.phi_pad),
                              _dp(self.phikk.view(np.float64)), dKE, dPE, dQI)

    def maxima(self, ref_range: bool = False):
        out = np.zeros(3)
        lib().oracle_maxima(ctypes.byref(self.cp), _dp(self.psi.view(np.float64)), _dp(self.phi_pad),
                            int(ref_range), _dp(out))
        return {"max_phi": out[0], "max_rho": out[1], "norm": out[2]}


class FastState:
    """The same FP64 step (poisson, kick, drift, Split_op_MPI, the two scans
    of Timestep_criteria) over an FFT library: scipy.fft (pocketfft, SIMD
    radix kernels) with `workers` threads for the 3-D transforms and the
    elementwise work split over x-planes in a thread pool.  A CPU baseline
    of the reference's algorithm on a tuned transform, beside the C
    restatement's plain radix-2 one; pinned against State (the C oracle) by
    tests/test_oracle_physics.py.  phi is held unpadded (N^3 doubles)."""

    def __init__(self, rp: RunParams, psi: np.ndarray, workers: int = 0):
        from concurrent.futures import ThreadPoolExecutor
        n = rp.n
        self.rp = rp
        self.psi = np.ascontiguousarray(psi, dtype=np.complex128).reshape(n, n, n).copy()
        self.phi = np.zeros((n, n, n))
        self.workers = workers or (os.cpu_count() or 1)
        self._pool = ThreadPoolExecutor(self.workers)
        self._kax = kvec(n, rp.xmax) ** 2                                  # pow(k, 2), struct3D.h:76-85
        self._mt = rp.mfdm / math.pow(c_light, 2)

    def close(self):
        self._pool.shutdown()

    def _planes(self, fn):
        n = self.rp.n
        step = max(1, -(-n // (4 * self.workers)))
        list(self._pool.map(lambda i0: fn(i0, min(n, i0 + step)), range(0, n, step)))

    def _ksq(self, i0, i1, nz):
        k = self._kax
        return k[i0:i1, None, None] + k[None, :, None] + k[None, None, :nz]

    def poisson(self):
        import scipy.fft as sf
        rp, n = self.rp, self.rp.n
        ms = rp.mfdm_solar

        def src(i0, i1):
            z = self.psi[i0:i1]
            self.phi[i0:i1] = ms * (z.real * z.real + z.imag * z.imag) - rp.rhoc      # :88
        self._planes(src)
        sk = sf.rfftn(self.phi, workers=self.workers)                                  # :94
        f = -4.0 * math.pi * bigG                                                       # :105
        h = n // 2 + 1

        def green(i0, i1):
            kk = self._ksq(i0, i1, h)
            with np.errstate(divide="ignore"):
                g = np.where(kk == 0, 0.0, f / kk)                                     # :101-108
            sk[i0:i1] *= g
        self._planes(green)
        self.phi = sf.irfftn(sk, s=(n, n, n), workers=self.workers)                    # :111-116 (norm 1/N^3)

    def kick(self, dPE: float, dQI: float):
        rp = self.rp
        cphi = dPE * self._mt / hbar
        crho = dQI * rp.g / hbar

        def k(i0, i1):
            z = self.psi[i0:i1]
            th = -0.5 * (cphi * self.phi[i0:i1] + crho * (z.real * z.real + z.imag * z.imag))   # :130-134
            self.psi[i0:i1] = z * (np.cos(th) + 1j * np.sin(th))
        self._planes(k)

    def drift(self, dKE: float):
        import scipy.fft as sf
        n = self.rp.n
        ke = 0.5 * hbar / self._mt                                                     # struct3D.h:121
        a = sf.fftn(self.psi, workers=self.workers, overwrite_x=True)                  # :136

        def d(i0, i1):
            th = -ke * dKE * self._ksq(i0, i1, n)
            a[i0:i1] *= np.cos(th) + 1j * np.sin(th)                                   # :141 (N^-3 in ifftn)
        self._planes(d)
        self.psi = sf.ifftn(a, workers=self.workers, overwrite_x=True)                 # :144

    def split_op(self, dKE: float, dPE: float, dQI: float):
        self.kick(dPE, dQI)
        self.drift(dKE)
        self.poisson()
        self.kick(dPE, dQI)

    def maxima(self):
        """max|phi| over every cell, max|psi|^2 and sum|psi|^2 (:26-27)."""
        n = self.rp.n
        part = {}

        def m(i0, i1):
            z = self.psi[i0:i1]
            rr = z.real * z.real + z.imag * z.imag
            part[i0] = (float(np.max(np.abs(self.phi[i0:i1]))), float(np.max(rr)), float(np.sum(rr)))
        self._planes(m)
        v = [part[k] for k in sorted(part)]
        return {"max_phi": max(x[0] for x in v), "max_rho": max(x[1] for x in v), "norm": sum(x[2] for x in v)}


# ----------------------------------------------------------- time step
def _Ha(x, n):
    """Ha{2,3,4}_preintegrat, src/struct3D.h:127-146"""
    H = H_0 * math.sqrt(Omega_m0 / math.pow(x, 3) + (1.0 - Omega_m0))
    return 1. / (H * math.pow(x, n))


def integral(a0: float, a1: float, n: int) -> float:
    """gsl_integration_qags(.., a0, a1, 0, 1e-10, 1000, ..) of 1/(H a^n)
    (src/CosmoBEC_MPI_stan.cpp:66-68) via QUADPACK's QAGS."""
    from scipy.integrate import quad
    val, _err = quad(_Ha, a0, a1, args=(n,), epsabs=0.0, epsrel=1e-10, limit=1000)
    return val


def timestep_criteria(rp: RunParams, a: float, max_phi: float, max_rho: float):
    """Timestep_criteria, src/CosmoBEC_MPI_stan.cpp:19-78.

    Returns (dKE, dPE, dQI, a_next, which) with which in {3, 2, 4} naming
    the limiting criterion as the reference's ``n``."""
    m_tilde = rp.mfdm / math.pow(c_light, 2)
    dt_KE = 4. / (3. * math.pi) * m_tilde / hbar * math.pow(rp.dx, 2) * math.pow(a, 2)   # :32
    dt_PE = 2. * math.pi * hbar / m_tilde / max_phi * a                                   # :38
    g = rp.g
    dt_QI = 2. * math.pi * hbar / abs(g) / max_rho * math.pow(a, 3) if g != 0 else 0.0     # :41-42
    max_dt, which = dt_KE, 3                                                              # :49-58
    if max_dt > dt_PE:
        max_dt, which = dt_PE, 2
    if max_dt > dt_QI and g != 0:
        max_dt, which = dt_QI, 4
    H = H_0 * math.sqrt(Omega_m0 / math.pow(a, 3) + (1.0 - Omega_m0))                     # :61
    da = H * a * max_dt
    a_next = a + da
    dKE = integral(a, a + da, 3)
    dPE = integral(a, a + da, 2)
    dQI = integral(a, a + da, 4)
    return dKE, dPE, dQI, a_next, which


def simulate(rp: RunParams, psi0: np.ndarray, a0: float, nsteps: int, ref_maxphi_range: bool = False,
             schedule=None):
    """The time loop of Simulation(), src/CosmoBEC_MPI_stan.cpp:212-304,
    without its snapshot triggers: poisson -> Timestep_criteria ->
    Split_op_MPI per step.  With ``schedule`` (list of (dKE, dPE, dQI)) the
    recorded time steps are replayed instead of recomputed.  Returns the
    final State and the list of (dKE, dPE, dQI, a_next) used."""
    st = State(rp, psi0)
    a = a0
    used = []
    for tn in range(nsteps):
        st.poisson()                                                   # :218
        if schedule is None:
            m = st.maxima(ref_maxphi_range)
            dKE, dPE, dQI, a, _ = timestep_criteria(rp, a, m["max_phi"], m["max_rho"])
        else:
            dKE, dPE, dQI = schedule[tn]
        used.append((dKE, dPE, dQI, a))
        st.split_op(dKE, dPE, dQI)                                     # :304
    return st, used


class Triggers:
    """The snapshot triggers of Simulation, src/CosmoBEC_MPI_stan.cpp:238-284
    (the _stan chain), evaluated after Timestep_criteria has advanced a:
    __call__ returns (snapshot number or None, break flag)."""

    def __init__(self):
        self.write_num, self.sten_counter, self.write_sten_counter = 0, 0, 2      # :208-210

    def __call__(self, a: float, tn: int, twrite: int, timesteps: int):
        t, stop = None, False
        if a > 1:                                                      # :238
            t, stop = 999999, True
            self.write_num += 1
        elif a > 0.714 and self.write_num == 4:                        # :244
            t = 999998
            self.write_num += 1
        elif a > 0.5 and self.write_num == 2:                          # :250
            t, stop = 999997, True
            self.write_num += 1
        elif a > 0.25 and self.write_num == 2:                         # :262
            t = self.write_sten_counter
            self.write_sten_counter += 1
        elif a > 0.25 and self.write_num == 1:                         # :269
            t = 999995
            self.write_num += 1
            self.sten_counter = 0
        elif a > 0.105 and self.write_num == 0:                        # :276
            t = 999994
            self.write_num += 1
        elif tn % twrite == 0 or tn == timesteps - 1:                  # :282
            t = tn + 1
        self.sten_counter += 1                                         # :287
        return t, stop


def run_main(rp: RunParams, psi0: np.ndarray, a0: float, twrite: int, timesteps: int, ti: int = 0):
    """The reference's run from a fresh IC (main + Simulation,
    src/CosmoBEC_MPI_stan.cpp:212-310): per step poisson (:218),
    Timestep_criteria (:219), the snapshot triggers (:238-284; a snapshot
    holds psi before the step and the advanced a, :219 then :276), then
    Split_op_MPI (:304).  Returns [(t, a, psi), ...] for every snapshot
    written, in order, and the number of steps run."""
    st = State(rp, psi0)
    a = a0
    trig = Triggers()
    snaps = []
    steps = 0
    for tn in range(ti, timesteps):
        st.poisson()
        m = st.maxima()
        dKE, dPE, dQI, a, _ = timestep_criteria(rp, a, m["max_phi"], m["max_rho"])
        t, stop = trig(a, tn, twrite, timesteps)
        if t is not None:
            snaps.append((t, a, st.psi.copy()))
            if stop:
                break
        st.split_op(dKE, dPE, dQI)
        steps += 1
    return snaps, steps


def ic_from_particles(rp: RunParams, chunks, box: float, pmass: float, a: float, ref_p1: bool = False,
                      f32: bool = False, grids: bool = False):
    """read_gadget after the file reads (src/gadget_IC.h:610-672): CIC of
    every chunk, mass-weighted velocities, velocity-divergence phase and
    psi.  box kpc/h, pmass Msun/h.  f32 reproduces the reference's float
    velocity grids.  Returns (psi, theta, skipped), and with grids=True
    also (rho, vx, vy, vz): the deposited density and the mass-weighted
    velocity grids (vx is what write_binary_vg writes, src/write3D.h:24-53)."""
    n = rp.n
    rho = np.zeros((n, n, n))
    vs = [np.zeros((n, n, n)) for _ in range(3)]
    fp = ctypes.POINTER(ctypes.c_float)
    skipped = 0
    for pos, vel in chunks:
        p = np.ascontiguousarray(pos, dtype=np.float32)
        v = np.ascontiguousarray(vel, dtype=np.float32)
        skipped += lib().oracle_cic_chunk(n, box, pmass, p.ctypes.data_as(fp), v.ctypes.data_as(fp), p.shape[0],
                                          int(ref_p1), int(f32), _dp(rho), _dp(vs[0]), _dp(vs[1]), _dp(vs[2]))
    psi = np.zeros((n, n, n), np.complex128)
    theta = np.zeros((n, n, n), np.complex128)
    cp = rp.c_params()
    lib().oracle_ic_assemble(ctypes.byref(cp), pmass, a, rp.dx, h_0, m2Mpc, s2Myr, int(f32), _dp(rho),
                             _dp(vs[0]), _dp(vs[1]), _dp(vs[2]), _dp(psi.view(np.float64)),
                             _dp(theta.view(np.float64)))
    if grids:
        return psi, theta, skipped, (rho, vs[0], vs[1], vs[2])
    return psi, theta, skipped


def header_rhoc(nchunks: int, npart_chunk: int, pmass: float, box: float) -> float:
    """par.rhoc from the Gadget header, src/gadget_IC.h:319 (pmass Msun/h)."""
    return nchunks * npart_chunk * pmass / math.pow(box, 3) * math.pow(1e3, 3) * math.pow(h_0, 2)


def set_threads(t: int):
    lib().oracle_set_threads(int(t))


def num_threads() -> int:
    return lib().oracle_num_threads()
