"""The C ABI: libifdm.so loads and exports every symbol include/ifdm.h
declares; argument validation happens before any device work.  CPU only
(no compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "ifdm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ifdm_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    import ifdm_code_amd as F
    L = F.lib()
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(F.EXPORTS) == sorted(syms)


def test_abi_version_and_constants():
    import ifdm_code_amd as F
    from oracle import oracle as O
    assert F.lib().ifdm_abi_version() == 2
    for name, val in [("hbar", O.hbar), ("c", O.c_light), ("bigG", O.bigG), ("H_0", O.H_0),
                      ("h_0", O.h_0), ("Omega_m0", O.Omega_m0), ("rhoc", O.rhoc_const),
                      ("cm2Mpc", O.cm2Mpc), ("kg2eV", O.kg2eV), ("kg2solar", O.kg2solar)]:
        assert F.constant(name) == val, name
    import math
    assert math.isnan(F.constant("nope"))


@pytest.mark.parametrize("n", [0, 7, 100, 2048])
def test_create_rejects_bad_sizes_without_touching_the_device(n):
    import ifdm_code_amd as F
    with pytest.raises(F.IfdmError, match="power of two"):
        F.Solver(F.Config(n=n, xmax=5.0))


def test_create_rejects_bad_physics():
    import ifdm_code_amd as F
    with pytest.raises(F.IfdmError, match="positive"):
        F.Solver(F.Config(n=16, xmax=-1.0))
