"""GPU: ifdm_power_spectrum_noalloc (bins in the library's fixed device
scratch, nothing allocated) against ifdm_power_spectrum (bins in a buffer
it allocates), through the C ABI, on two calls in a row, the second on a
smaller mesh with fewer bins: the scratch keeps the first call's contents,
so stale bins would show in the second."""
import ctypes

import numpy as np
import pytest

from tests.fields import make_psi

pytestmark = pytest.mark.gpu


def _pk(F, s, nb, noalloc):
    k, pk, nm = np.zeros(nb), np.zeros(nb), np.zeros(nb)
    fn = "ifdm_power_spectrum_noalloc" if noalloc else "ifdm_power_spectrum"
    rc = getattr(F.lib(), fn)(s._h, nb, k.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                              pk.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                              nm.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0, F.lib().ifdm_last_error(s._h).decode()
    return k, pk, nm


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_power_spectrum_matches_allocating_one(prec):
    import ifdm_code_amd as F
    cases = [(64, 40, 3), (32, 12, 4)]          # (mesh, bins, seed): the second call is smaller
    for n, nb, seed in cases:
        print(f"[noalloc P(k)] precision={prec} n={n} nbins={nb}", flush=True)
        cfg = F.Config.from_cli(2.5e-24, 0.0, 5.0, n, precision=prec)
        with F.Solver(cfg) as s:
            s.set_psi(make_psi(n, seed, cfg.mfdm_solar))
            # the first call sets the mixed mode's mean density from the
            # device reduction, as every later call leaves it
            _pk(F, s, nb, False)
            k0, pk0, nm0 = _pk(F, s, nb, False)
            st0 = s.poisson()
            k1, pk1, nm1 = _pk(F, s, nb, True)
            st1 = s.poisson()
        assert np.array_equal(nm1, nm0)                       # integer-valued mode counts
        np.testing.assert_allclose(k1, k0, rtol=1e-13, atol=0)
        np.testing.assert_allclose(pk1, pk0, rtol=1e-12, atol=0)
        assert st1 == st0
        assert (pk1 > 0).all() and nm1.sum() > 0


def test_noalloc_power_spectrum_rejects_too_many_bins():
    import ifdm_code_amd as F
    cfg = F.Config.from_cli(2.5e-24, 0.0, 5.0, 16)
    with F.Solver(cfg) as s:
        s.set_psi(make_psi(16, 1, cfg.mfdm_solar))
        with pytest.raises(F.IfdmError, match="nbins must be in 1..4096"):
            s.power_spectrum(4097, noalloc=True)
        k, pk, nm = s.power_spectrum(8, noalloc=True)
        assert (nm > 0).all()
