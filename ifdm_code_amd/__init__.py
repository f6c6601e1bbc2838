"""ifdm_code_amd -- MI355X-native Schrodinger-Poisson kick-drift-kick solver.

Python binding (ctypes) of libifdm.so, the C ABI declared in include/ifdm.h.
The names mirror the reference's free functions (jchan192/IFDM_Code,
src/CosmoBEC_MPI_stan.cpp): ``poisson`` (:80), ``split_op`` (Split_op_MPI,
:121), ``timestep_criteria`` (:19).  All compute runs in the HIP library;
this module only moves host buffers and scalars across the boundary, and it
raises if the library is missing or a call fails.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# IFDM_LIBRARY overrides the library path (e.g. to compare two builds)
LIB_PATH = os.environ.get("IFDM_LIBRARY") or os.path.join(_HERE, "libifdm.so")
_LIB = None

FP64 = 0
MIXED = 1
# passes of one step as timed by ifdm_get_profile / ifdm_time_passes
# (include/ifdm.h, IFDM_NPASS)
NPASS = 9


class IfdmError(RuntimeError):
    pass


class _Config(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("precision", ctypes.c_int), ("xmax", ctypes.c_double),
                ("mfdm", ctypes.c_double), ("g", ctypes.c_double), ("rhoc", ctypes.c_double),
                ("ref_maxphi_range", ctypes.c_int), ("device", ctypes.c_int), ("plane_pad", ctypes.c_int)]


class _Stats(ctypes.Structure):
    _fields_ = [("max_phi", ctypes.c_double), ("max_rho", ctypes.c_double), ("norm", ctypes.c_double)]

    def as_dict(self):
        return {"max_phi": self.max_phi, "max_rho": self.max_rho, "norm": self.norm}


class _Dt(ctypes.Structure):
    _fields_ = [("dKE", ctypes.c_double), ("dPE", ctypes.c_double), ("dQI", ctypes.c_double)]


class _Trig(ctypes.Structure):
    _fields_ = [("write_num", ctypes.c_int), ("sten_counter", ctypes.c_int), ("write_sten_counter", ctypes.c_int)]


class OutputTriggers:
    """The snapshot trigger chain of the reference's time loop
    (src/CosmoBEC_MPI_stan.cpp:238-284; variant 1 = src/CosmoBEC_MPI.cpp)."""

    def __init__(self, variant: int = 0, a_triggers: bool = True):
        self.variant = variant
        self.a_triggers = a_triggers
        self.st = _Trig(0, 0, 2)

    def __call__(self, a: float, tn: int, twrite: int, t_stop: int):
        """(snapshot number or None, stop flag) for this step."""
        t, stop = ctypes.c_int(), ctypes.c_int()
        r = lib().ifdm_output_trigger(self.variant, int(self.a_triggers), a, tn, twrite, t_stop,
                                      ctypes.byref(self.st), ctypes.byref(t), ctypes.byref(stop))
        if r < 0:
            raise IfdmError("ifdm_output_trigger: bad arguments")
        return (t.value if r == 1 else None), bool(stop.value)


class _ICInfo(ctypes.Structure):
    _fields_ = [("box", ctypes.c_double), ("pmass", ctypes.c_double), ("a", ctypes.c_double),
                ("rhoc", ctypes.c_double), ("npart", ctypes.c_longlong), ("skipped", ctypes.c_longlong)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


EXPORTS = [
    "ifdm_power_spectrum", "ifdm_power_spectrum_noalloc",
    "ifdm_run_checkpointed", "ifdm_adjoint", "ifdm_adjoint_device", "ifdm_set_psi_device", "ifdm_get_psi_device",
    "ifdm_read_gadget", "ifdm_ic_begin", "ifdm_ic_deposit", "ifdm_ic_finish", "ifdm_set_rhoc",
    "ifdm_write_snapshot", "ifdm_read_snapshot",
    "ifdm_create", "ifdm_destroy", "ifdm_last_error", "ifdm_create_error", "ifdm_set_psi", "ifdm_get_psi",
    "ifdm_get_phi", "ifdm_set_phi", "ifdm_poisson", "ifdm_kick", "ifdm_kdk_step", "ifdm_step", "ifdm_flush", "ifdm_run", "ifdm_drift",
    "ifdm_fft3d", "ifdm_output_trigger", "ifdm_simulate", "ifdm_simulate_ex", "ifdm_time_passes",
    "ifdm_set_profiling", "ifdm_get_profile", "ifdm_stream_copy", "ifdm_set_option", "ifdm_timestep_criteria", "ifdm_scale_integral", "ifdm_constant", "ifdm_abi_version",
]


def lib():
    """Load libifdm.so; raises IfdmError when it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise IfdmError(f"{LIB_PATH} is missing: build it with `python -m ifdm_code_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int
    L.ifdm_create.argtypes = [ctypes.POINTER(_Config), ctypes.POINTER(vp)]
    L.ifdm_destroy.argtypes = [vp]
    L.ifdm_destroy.restype = None
    L.ifdm_last_error.argtypes = [vp]
    L.ifdm_last_error.restype = ctypes.c_char_p
    L.ifdm_create_error.restype = ctypes.c_char_p
    for f in ("ifdm_set_psi", "ifdm_get_psi", "ifdm_get_phi", "ifdm_set_phi"):
        getattr(L, f).argtypes = [vp, dp]
    L.ifdm_poisson.argtypes = [vp, ctypes.POINTER(_Stats)]
    L.ifdm_kick.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    L.ifdm_kdk_step.argtypes = [vp, ctypes.POINTER(_Dt), ctypes.POINTER(_Stats)]
    L.ifdm_run.argtypes = [vp, ctypes.POINTER(_Dt), ip, ctypes.POINTER(_Stats)]
    L.ifdm_step.argtypes = [vp, ctypes.POINTER(_Dt), ctypes.POINTER(_Stats)]
    L.ifdm_flush.argtypes = [vp]
    L.ifdm_drift.argtypes = [vp, ctypes.c_double]
    L.ifdm_simulate.argtypes = [vp, ctypes.c_double, ip, dp, ctypes.POINTER(_Dt)]
    L.ifdm_simulate_ex.argtypes = [vp, ctypes.c_double, ip, dp, ctypes.POINTER(_Dt), ip]
    L.ifdm_time_passes.argtypes = [vp, ctypes.POINTER(_Dt), ip, ctypes.POINTER(ctypes.c_float)]
    L.ifdm_fft3d.argtypes = [vp, ip]
    L.ifdm_power_spectrum.argtypes = [vp, ip, dp, dp, dp]
    L.ifdm_power_spectrum_noalloc.argtypes = [vp, ip, dp, dp, dp]
    L.ifdm_run_checkpointed.argtypes = [vp, ctypes.POINTER(_Dt), ip, ip]
    L.ifdm_adjoint.argtypes = [vp, ctypes.POINTER(_Dt), ip, dp]
    L.ifdm_adjoint_device.argtypes = [vp, ctypes.POINTER(_Dt), ip, vp]
    L.ifdm_set_psi_device.argtypes = [vp, vp]
    L.ifdm_get_psi_device.argtypes = [vp, vp]
    fp = ctypes.POINTER(ctypes.c_float)
    ll = ctypes.c_longlong
    L.ifdm_read_gadget.argtypes = [vp, ctypes.c_char_p, ip, ip, ctypes.c_char_p, ctypes.POINTER(_ICInfo)]
    L.ifdm_ic_begin.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    L.ifdm_ic_deposit.argtypes = [vp, fp, fp, ll, ip]
    L.ifdm_ic_finish.argtypes = [vp, ctypes.c_double, dp, fp, ctypes.POINTER(ll)]
    L.ifdm_set_rhoc.argtypes = [vp, ctypes.c_double]
    L.ifdm_write_snapshot.argtypes = [vp, ctypes.c_char_p, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    L.ifdm_read_snapshot.argtypes = [vp, ctypes.c_char_p, dp, dp, dp]
    L.ifdm_set_profiling.argtypes = [vp, ip]
    L.ifdm_stream_copy.argtypes = [vp, ctypes.c_size_t, ip, dp]
    L.ifdm_set_option.argtypes = [vp, ctypes.c_char_p, ctypes.c_long]
    L.ifdm_get_profile.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_long)]
    L.ifdm_timestep_criteria.argtypes = [ctypes.POINTER(_Config), ctypes.c_double, ctypes.POINTER(_Stats),
                                         ctypes.POINTER(_Dt), dp, ctypes.POINTER(ctypes.c_int)]
    L.ifdm_output_trigger.argtypes = [ip, ip, ctypes.c_double, ip, ip, ip, ctypes.POINTER(_Trig),
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.ifdm_scale_integral.argtypes = [ctypes.c_double, ctypes.c_double, ip]
    L.ifdm_scale_integral.restype = ctypes.c_double
    L.ifdm_constant.argtypes = [ctypes.c_char_p]
    L.ifdm_constant.restype = ctypes.c_double
    _LIB = L
    return L


def constant(name: str) -> float:
    """A constant of src/Cosmoinput3D.h:2-17 as the library holds it."""
    return lib().ifdm_constant(name.encode())


@dataclass
class Config:
    """Run configuration.  ``from_cli`` applies the reference's argv unit
    conversions (src/CosmoBEC_MPI_stan.cpp:329-332)."""
    n: int
    xmax: float            # Mpc (half box)
    mfdm: float = 2.5e-24  # eV
    g: float = 0.0         # eV Mpc^3
    rhoc: float = 3.51059e10
    precision: int = FP64
    ref_maxphi_range: bool = False
    device: int = 0
    plane_pad: int = 0     # device cells after each x-plane (0: built-in, <0: none)

    @classmethod
    def from_cli(cls, mfdm_ev: float, g_ev_cm3: float, xmax_mpc_h: float, res: int, **kw):
        return cls(n=res, xmax=xmax_mpc_h / constant("h_0"), mfdm=mfdm_ev,
                   g=g_ev_cm3 / constant("cm2Mpc") ** 3, **kw)

    @property
    def mfdm_solar(self) -> float:
        return self.mfdm * constant("kg2eV") / constant("kg2solar")

    def _c(self) -> _Config:
        return _Config(self.n, self.precision, self.xmax, self.mfdm, self.g, self.rhoc,
                       int(self.ref_maxphi_range), self.device, int(self.plane_pad))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def timestep_criteria(cfg: Config, a: float, stats: dict):
    """Timestep_criteria (src/CosmoBEC_MPI_stan.cpp:19-78), host C++ in the
    library.  Returns ((dKE, dPE, dQI), a_next, which)."""
    st = _Stats(stats["max_phi"], stats["max_rho"], stats.get("norm", 0.0))
    dt = _Dt()
    an = ctypes.c_double()
    which = ctypes.c_int()
    rc = lib().ifdm_timestep_criteria(ctypes.byref(cfg._c()), a, ctypes.byref(st), ctypes.byref(dt),
                                      ctypes.byref(an), ctypes.byref(which))
    if rc:
        raise IfdmError(f"ifdm_timestep_criteria failed ({rc})")
    return (dt.dKE, dt.dPE, dt.dQI), an.value, which.value


def scale_integral(a0: float, a1: float, power: int) -> float:
    return lib().ifdm_scale_integral(a0, a1, power)


class Solver:
    """A device context holding psi, phi and the Poisson half spectrum."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self._h = ctypes.c_void_p()
        rc = lib().ifdm_create(ctypes.byref(cfg._c()), ctypes.byref(self._h))
        if rc:
            raise IfdmError(f"ifdm_create failed ({rc}): {lib().ifdm_create_error().decode()}")

    def close(self):
        if self._h:
            lib().ifdm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc: int, what: str):
        if rc:
            raise IfdmError(f"{what} failed ({rc}): {lib().ifdm_last_error(self._h).decode()}")

    # ---- data movement
    def set_psi(self, psi: np.ndarray):
        n = self.cfg.n
        a = np.ascontiguousarray(psi, dtype=np.complex128)
        if a.size != n ** 3:
            raise ValueError(f"psi must hold {n}^3 values")
        self._chk(lib().ifdm_set_psi(self._h, _dp(a.view(np.float64))), "ifdm_set_psi")

    def get_psi(self) -> np.ndarray:
        n = self.cfg.n
        out = np.empty((n, n, n), np.complex128)
        self._chk(lib().ifdm_get_psi(self._h, _dp(out.view(np.float64))), "ifdm_get_psi")
        return out

    def get_phi(self) -> np.ndarray:
        n = self.cfg.n
        out = np.empty((n, n, n), np.float64)
        self._chk(lib().ifdm_get_phi(self._h, _dp(out)), "ifdm_get_phi")
        return out

    def set_phi(self, phi: np.ndarray):
        a = np.ascontiguousarray(phi, dtype=np.float64)
        if a.size != self.cfg.n ** 3:
            raise ValueError("phi size mismatch")
        self._chk(lib().ifdm_set_phi(self._h, _dp(a)), "ifdm_set_phi")

    # ---- diagnostics
    def power_spectrum(self, nbins: int | None = None, noalloc: bool = False):
        """(k [1/Mpc], P(k) [Mpc^3], modes) on shells of width 2 pi / L.
        noalloc: the entry point that allocates no device memory
        (ifdm_power_spectrum_noalloc, at most 4096 bins)."""
        nb = nbins or self.cfg.n // 2
        k, pk, nm = np.zeros(nb), np.zeros(nb), np.zeros(nb)
        fn = "ifdm_power_spectrum_noalloc" if noalloc else "ifdm_power_spectrum"
        self._chk(getattr(lib(), fn)(self._h, nb, _dp(k), _dp(pk), _dp(nm)), fn)
        return k, pk, nm

    # ---- adjoint
    @staticmethod
    def _sched(schedule):
        m = len(schedule)
        return (_Dt * max(m, 1))(*[_Dt(*x) for x in schedule]), m

    def run_checkpointed(self, schedule, K: int) -> int:
        """Forward run from the current psi (phi re-solved on entry) keeping
        checkpoints for one reverse sweep.  Returns a serial number that
        changes with every such run (the autograd binding checks it)."""
        arr, m = self._sched(schedule)
        self._chk(lib().ifdm_run_checkpointed(self._h, arr, m, int(K)), "ifdm_run_checkpointed")
        self._ck_runs = getattr(self, "_ck_runs", 0) + 1
        self.ck_serial = self._ck_runs
        return self.ck_serial

    def adjoint(self, schedule, g_T: np.ndarray) -> np.ndarray:
        """Gradient w.r.t. psi_0 from the gradient w.r.t. psi_T (both in the
        dL/dRe + i dL/dIm convention) after run_checkpointed(schedule)."""
        arr, m = self._sched(schedule)
        g = np.ascontiguousarray(g_T, dtype=np.complex128).copy()
        self._chk(lib().ifdm_adjoint(self._h, arr, m, _dp(g.view(np.float64))), "ifdm_adjoint")
        return g

    def adjoint_device(self, schedule, g_ptr: int):
        arr, m = self._sched(schedule)
        self._chk(lib().ifdm_adjoint_device(self._h, arr, m, ctypes.c_void_p(g_ptr)), "ifdm_adjoint_device")

    def set_psi_device(self, ptr: int):
        self._chk(lib().ifdm_set_psi_device(self._h, ctypes.c_void_p(ptr)), "ifdm_set_psi_device")

    def get_psi_device(self, ptr: int):
        self._chk(lib().ifdm_get_psi_device(self._h, ctypes.c_void_p(ptr)), "ifdm_get_psi_device")

    # ---- initial conditions and snapshots
    def read_gadget(self, icpath: str, nchunks: int = 4, ref_ghosts: bool = False, voutput: str | None = None):
        info = _ICInfo()
        self._chk(lib().ifdm_read_gadget(self._h, icpath.encode(), nchunks, int(ref_ghosts),
                                         voutput.encode() if voutput else None, ctypes.byref(info)),
                  "ifdm_read_gadget")
        return info.as_dict()

    def ic_from_particles(self, chunks, box: float, pmass: float, a: float, ref_ghosts: bool = False,
                          want_theta: bool = False, want_vx: bool = False):
        """chunks: iterable of (pos, vel) float32 arrays of shape (np, 3).
        Returns (theta or None, skipped), and with want_vx also the
        mass-weighted float vx grid (what voutput.dat holds).  The deposit
        follows the "ic_deposit" option (set_option): 0, the default, the
        reference's order and float velocity grids; 1, FP64 atomics."""
        self._chk(lib().ifdm_ic_begin(self._h, box, pmass), "ifdm_ic_begin")
        fp = ctypes.POINTER(ctypes.c_float)
        for pos, vel in chunks:
            p = np.ascontiguousarray(pos, dtype=np.float32)
            v = np.ascontiguousarray(vel, dtype=np.float32)
            self._chk(lib().ifdm_ic_deposit(self._h, p.ctypes.data_as(fp), v.ctypes.data_as(fp), p.shape[0],
                                            int(ref_ghosts)), "ifdm_ic_deposit")
        n = self.cfg.n
        theta = np.empty((n, n, n), np.complex128) if want_theta else None
        vx = np.empty((n, n, n), np.float32) if want_vx else None
        sk = ctypes.c_longlong()
        self._chk(lib().ifdm_ic_finish(self._h, a, _dp(theta.view(np.float64)) if want_theta else None,
                                       vx.ctypes.data_as(fp) if want_vx else None, ctypes.byref(sk)),
                  "ifdm_ic_finish")
        if want_vx:
            return theta, sk.value, vx
        return theta, sk.value

    def set_rhoc(self, rhoc: float):
        self._chk(lib().ifdm_set_rhoc(self._h, rhoc), "ifdm_set_rhoc")

    def write_snapshot(self, path: str, a: float, mfdm_ev: float, g_ev_cm3: float):
        self._chk(lib().ifdm_write_snapshot(self._h, path.encode(), a, mfdm_ev, g_ev_cm3), "ifdm_write_snapshot")

    def read_snapshot(self, path: str):
        a, m, g = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._chk(lib().ifdm_read_snapshot(self._h, path.encode(), ctypes.byref(a), ctypes.byref(m),
                                           ctypes.byref(g)), "ifdm_read_snapshot")
        return a.value, m.value, g.value

    # ---- the hot path
    def poisson(self) -> dict:
        st = _Stats()
        self._chk(lib().ifdm_poisson(self._h, ctypes.byref(st)), "ifdm_poisson")
        return st.as_dict()

    def kick(self, dPE: float, dQI: float):
        self._chk(lib().ifdm_kick(self._h, dPE, dQI), "ifdm_kick")

    def drift(self, dKE: float):
        self._chk(lib().ifdm_drift(self._h, dKE), "ifdm_drift")

    def fft3d(self, sign: int):
        self._chk(lib().ifdm_fft3d(self._h, int(sign)), "ifdm_fft3d")

    def split_op(self, dKE: float, dPE: float, dQI: float) -> dict:
        st = _Stats()
        dt = _Dt(dKE, dPE, dQI)
        self._chk(lib().ifdm_kdk_step(self._h, ctypes.byref(dt), ctypes.byref(st)), "ifdm_kdk_step")
        return st.as_dict()

    def step(self, dKE: float, dPE: float, dQI: float) -> dict:
        """One step with the second half-kick deferred (ifdm_step)."""
        st = _Stats()
        self._chk(lib().ifdm_step(self._h, ctypes.byref(_Dt(dKE, dPE, dQI)), ctypes.byref(st)), "ifdm_step")
        return st.as_dict()

    def flush(self):
        self._chk(lib().ifdm_flush(self._h), "ifdm_flush")

    def run(self, schedule, stats: bool = False):
        """Replay a schedule [(dKE, dPE, dQI), ...] with merged half-kicks."""
        m = len(schedule)
        arr = (_Dt * max(m, 1))(*[_Dt(*s) for s in schedule])
        sts = (_Stats * max(m, 1))() if stats else None
        self._chk(lib().ifdm_run(self._h, arr, m, sts), "ifdm_run")
        return [s.as_dict() for s in sts[:m]] if stats else None

    def simulate_native(self, a0: float, nsteps: int, resume: bool = False, keep_pending: bool = False):
        """ifdm_simulate (ifdm_simulate_ex with resume / keep_pending): the
        adaptive loop in the library with merged half-kicks.  Returns
        (a_final, [(dKE, dPE, dQI), ...])."""
        sched = (_Dt * max(nsteps, 1))()
        a = ctypes.c_double()
        flags = (1 if resume else 0) | (2 if keep_pending else 0)
        self._chk(lib().ifdm_simulate_ex(self._h, a0, nsteps, ctypes.byref(a), sched, flags), "ifdm_simulate_ex")
        return a.value, [(d.dKE, d.dPE, d.dQI) for d in sched[:nsteps]]

    def stream_copy(self, nbytes: int, iters: int = 10) -> float:
        """GB/s of a plain device copy of nbytes (read + write counted)."""
        g = ctypes.c_double()
        self._chk(lib().ifdm_stream_copy(self._h, nbytes, iters, ctypes.byref(g)), "ifdm_stream_copy")
        return g.value

    def set_option(self, name: str, value: int):
        self._chk(lib().ifdm_set_option(self._h, name.encode(), int(value)), "ifdm_set_option")

    def set_profiling(self, on: bool):
        self._chk(lib().ifdm_set_profiling(self._h, int(on)), "ifdm_set_profiling")

    def get_profile(self):
        """(summed ms per pass: drift table, A, B, C, F, G, H, I, reduction;
        number of profiled steps)"""
        ms = (ctypes.c_double * NPASS)()
        cnt = ctypes.c_long()
        self._chk(lib().ifdm_get_profile(self._h, ms, ctypes.byref(cnt)), "ifdm_get_profile")
        return [float(x) for x in ms], cnt.value

    def time_passes(self, dt, iters: int = 3):
        """Average ms of each pass (drift table, A, B, C, F, G, H, I,
        reduction) of one step."""
        out = (ctypes.c_float * NPASS)()
        self._chk(lib().ifdm_time_passes(self._h, ctypes.byref(_Dt(*dt)), iters, out), "ifdm_time_passes")
        return [float(x) for x in out]

    def simulate(self, a0: float, nsteps: int):
        """Adaptive run with the reference's loop (src/CosmoBEC_MPI_stan.cpp:
        212-304) minus snapshot I/O: poisson -> Timestep_criteria -> step,
        one unmerged Split_op_MPI per step."""
        st = self.poisson()
        a = a0
        used = []
        for _ in range(nsteps):
            dt, a, _w = timestep_criteria(self.cfg, a, st)
            used.append((*dt, a))
            st = self.split_op(*dt)
        return used
