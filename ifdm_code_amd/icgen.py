"""Synthetic cosmological initial conditions in the reference's input format.

The reference reads particle ICs produced by MUSIC and an external sorter
(README.md:1, src/gadget_IC.h:259-353): four Gadget-1 files <icpath>.0..3,
each sorted by x.  Neither tool ships with it, so this module makes the same
kind of input:

* N^3 particles on the cell-centred lattice q = (i + 1/2) Lg;
* Zel'dovich displacements s with div s = -delta, s_k = i k delta_k / k^2,
  from a seeded Gaussian field with P(k) ~ k^slope for 0 < k <= k_Nyquist,
  scaled to a chosen rms |s| (in cells);
* Gadget velocities u = v / sqrt(a), v = a H(a) f(a) s, f = Omega_m(a)^0.55;
* particle mass chosen so that the reader's mean density
  (src/gadget_IC.h:319) is the reference's rho_c = 3.51059e10 Msun/Mpc^3.

Units: positions kpc/h, velocities km/s (Gadget internal), masses
1e10 Msun/h.  Cosmology: h = 0.677, Omega_m = 0.276 (src/Cosmoinput3D.h).
"""
from __future__ import annotations

import math
import os
import struct

import numpy as np

H0_LITTLE = 0.677
OMEGA_M = 0.276
RHOC = 3.51059e10


def growth_rate(a: float) -> float:
    e2 = OMEGA_M / a ** 3 + (1 - OMEGA_M)
    return (OMEGA_M / a ** 3 / e2) ** 0.55


def zeldovich(n: int, box_mpc_h: float, z: float = 50.0, seed: int = 1234, slope: float = -2.5,
              rms_cells: float = 0.2, dtype=None, progress=None):
    """Returns (pos, vel, pmass, a): pos and vel float32 arrays (n^3, 3) in
    kpc/h and Gadget km/s, the particle mass in 1e10 Msun/h, and a.  The
    field is built in float64 up to 256^3 and in float32 above (memory:
    about 10 x 4 n^3 bytes at 1024^3).  progress(str), if given, is called
    after each transform (minutes each at 1024^3)."""
    say = progress or (lambda _m: None)
    if dtype is None:
        dtype = np.float64 if n <= 256 else np.float32
    a = 1.0 / (1.0 + z)
    box = box_mpc_h * 1e3                       # kpc/h
    lg = box / n
    rng = np.random.default_rng(seed)
    k1 = (np.fft.fftfreq(n) * n).astype(dtype)  # integer wave numbers
    kz = np.arange(n // 2 + 1, dtype=dtype)
    kk = k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kz[None, None, :] ** 2
    nz = (kk > 0) & (kk <= (n / 2) ** 2)
    amp = np.zeros_like(kk)
    amp[nz] = kk[nz] ** dtype(slope / 4)
    del nz
    dk = np.fft.rfftn(rng.standard_normal((n, n, n), dtype=dtype)) * amp
    del amp
    say("zeldovich: density field transformed")
    kk[0, 0, 0] = 1.0
    dk /= kk                                    # delta_k / k^2
    del kk
    # s_k = i k delta_k / k^2, one axis at a time
    s = np.empty((n ** 3, 3), np.float32)
    for ax in range(3):
        kc = (k1[:, None, None], k1[None, :, None], kz[None, None, :])[ax]
        s[:, ax] = np.fft.irfftn(1j * kc * dk, s=(n, n, n), axes=(0, 1, 2)).ravel()
        say(f"zeldovich: displacement {ax} done")
    del dk
    rms = math.sqrt(float(np.mean(np.sum(s.astype(np.float64) ** 2, axis=1))))
    s *= np.float32(rms_cells * lg / rms)       # kpc/h
    e = math.sqrt(OMEGA_M / a ** 3 + 1 - OMEGA_M)
    # v = a H f s: 100 h E(a) km/s/Mpc x s[kpc/h] / (1e3 h); Gadget u = v / sqrt(a)
    vel = (s * np.float32(a * 0.1 * e * growth_rate(a) / math.sqrt(a))).astype(np.float32)
    idx = (np.arange(n, dtype=np.float64) + 0.5) * lg
    pos = s.reshape(n, n, n, 3)
    pos[..., 0] += idx[:, None, None].astype(np.float32)
    pos[..., 1] += idx[None, :, None].astype(np.float32)
    pos[..., 2] += idx[None, None, :].astype(np.float32)
    pos = pos.reshape(-1, 3)
    np.mod(pos, np.float32(box), out=pos)
    pos[pos >= box] = 0.0                       # float rounding at the upper edge
    pmass = RHOC * box ** 3 / (n ** 3 * 1e10 * 1e9 * H0_LITTLE ** 2)
    return pos, vel, pmass, a


def _header(np1: int, ntot: int, pmass: float, a: float, box: float, nfiles: int) -> bytes:
    h = struct.pack("<6i", 0, np1, 0, 0, 0, 0)
    h += struct.pack("<6d", 0.0, pmass, 0.0, 0.0, 0.0, 0.0)
    h += struct.pack("<2d", a, 1.0 / a - 1.0)
    h += struct.pack("<2i", 0, 0)
    h += struct.pack("<6i", 0, ntot, 0, 0, 0, 0)
    h += struct.pack("<2i", 0, nfiles)
    h += struct.pack("<4d", box, OMEGA_M, 1 - OMEGA_M, H0_LITTLE)
    assert len(h) == 160
    return h + b"\0" * 96


def write_gadget(icpath: str, pos: np.ndarray, vel: np.ndarray, pmass: float, a: float, box_kpc_h: float,
                 nchunks: int = 4):
    """Sorts the particles by x and writes nchunks Gadget-1 files
    <icpath>.0 .. with equal particle counts (the 'sorted IC' layout)."""
    order = np.argsort(pos[:, 0], kind="stable")
    pos = pos[order]
    vel = vel[order]
    ntot = pos.shape[0]
    assert ntot % nchunks == 0
    per = ntot // nchunks
    d = os.path.dirname(icpath)
    if d:
        os.makedirs(d, exist_ok=True)
    for c in range(nchunks):
        p = np.ascontiguousarray(pos[c * per:(c + 1) * per], dtype="<f4")
        v = np.ascontiguousarray(vel[c * per:(c + 1) * per], dtype="<f4")
        ids = np.arange(c * per, (c + 1) * per, dtype="<u4")
        with open(f"{icpath}.{c}", "wb") as f:
            for blk in (_header(per, ntot, pmass, a, box_kpc_h, nchunks), p.tobytes(), v.tobytes(), ids.tobytes()):
                f.write(struct.pack("<i", len(blk)))
                f.write(blk)
                f.write(struct.pack("<i", len(blk)))
    return [(np.asarray(pos[c * per:(c + 1) * per]), np.asarray(vel[c * per:(c + 1) * per])) for c in range(nchunks)]


def read_gadget(icpath: str, nchunks: int = 4):
    """(chunks [(pos, vel), ...], header dict) from <icpath>.0 .. (numpy)."""
    chunks = []
    hdr = None
    for c in range(nchunks):
        with open(f"{icpath}.{c}", "rb") as f:
            buf = f.read()
        np1 = struct.unpack_from("<6i", buf, 4)[1]
        mass = struct.unpack_from("<6d", buf, 28)[1]
        a = struct.unpack_from("<d", buf, 4 + 72)[0]
        box = struct.unpack_from("<d", buf, 4 + 128)[0]
        off = 4 + 256 + 4 + 4
        pos = np.frombuffer(buf, "<f4", 3 * np1, off).reshape(np1, 3)
        off += 12 * np1 + 8
        vel = np.frombuffer(buf, "<f4", 3 * np1, off).reshape(np1, 3)
        chunks.append((pos.copy(), vel.copy()))
        hdr = {"npart": np1, "mass": mass, "a": a, "box": box}
    return chunks, hdr
