"""The phase sin/cos of the kernels (ifdm_code_amd/csrc/fastmath.h),
compiled as plain C++ on the CPU and checked against long-double libm."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SHIM = r'''
#include "fastmath.h"
extern "C" void t_expmi64(const double *th, double *re, double *im, long n) {
    for (long i = 0; i < n; ++i) ifdm::expmi_parts<double>(th[i], &re[i], &im[i]);
}
extern "C" void t_expmi32(const double *th, float *re, float *im, long n) {
    for (long i = 0; i < n; ++i) ifdm::expmi_parts<float>(th[i], &re[i], &im[i]);
}
'''


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    d = tmp_path_factory.mktemp("fastmath")
    src = d / "shim.cpp"
    src.write_text(SHIM)
    so = d / "shim.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "ifdm_code_amd", "csrc"), str(src), "-o", str(so)], check=True)
    L = ctypes.CDLL(str(so))
    return L


def _run(L, fn, th, dt):
    re = np.empty(th.size, dt)
    im = np.empty(th.size, dt)
    getattr(L, fn)(th.ctypes.data_as(ctypes.c_void_p), re.ctypes.data_as(ctypes.c_void_p),
                   im.ctypes.data_as(ctypes.c_void_p), ctypes.c_long(th.size))
    return re, im


def _reference(th):
    t = th.astype(np.longdouble)
    return np.cos(t), -np.sin(t)


@pytest.mark.parametrize("scale", [1e-3, 1.0, 7.0, 60.0, 1e4])
def test_fp64_phase(shim, scale):
    rng = np.random.default_rng(1)
    th = rng.uniform(-scale, scale, 200000)
    th[:4] = [0.0, np.pi / 2, -np.pi / 4, 3 * np.pi]
    re, im = _run(shim, "t_expmi64", th, np.float64)
    cr, ci = _reference(th)
    err = max(np.abs(re - cr).max(), np.abs(im - ci).max())
    assert float(err) < 3.5e-16


@pytest.mark.parametrize("scale", [1.0, 7.0, 1e4])
def test_fp32_phase_after_fp64_reduction(shim, scale):
    rng = np.random.default_rng(2)
    th = rng.uniform(-scale, scale, 200000)
    re, im = _run(shim, "t_expmi32", th, np.float32)
    cr, ci = _reference(th)
    err = max(np.abs(re - cr).max(), np.abs(im - ci).max())
    assert float(err) < 1.5e-7       # about 2 FP32 ulp of 1, independent of |theta|
