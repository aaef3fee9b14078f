"""Accuracy of the series forms in merzbild.jl_amd/csrc/mz_math.h (used on the
device by the collision and wall kernels), compiled for the host from the
same header and measured in ulps against the C library through numpy."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def mlib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("mzmath") / "libmzmath.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                    os.path.join(HERE, "native", "math_check.cpp"), "-o", out], check=True)
    L = C.CDLL(out)
    P = C.POINTER(C.c_double)
    L.t_log.argtypes = [P, P, C.c_long]
    L.t_pow.argtypes = [P, C.c_double, P, C.c_long]
    L.t_sincos.argtypes = [P, P, P, C.c_long]
    return L


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ulp_err(got, ref):
    return np.abs(got - ref) / np.spacing(np.abs(ref))


def test_log_within_2_ulp(mlib):
    rng = np.random.default_rng(1)
    x = np.concatenate([10.0 ** rng.uniform(-50, 8, 400000), rng.uniform(0.5, 2.0, 200000),
                        1.0 + rng.uniform(-1e-6, 1e-6, 10000), [1.0, 2.0, 0.5, 1e-50, np.sqrt(0.5)]])
    y = np.empty_like(x)
    mlib.t_log(_p(x), _p(y), len(x))
    ref = np.log(x)
    nz = ref != 0
    assert np.all(y[~nz] == 0.0)
    assert _ulp_err(y[nz], ref[nz]).max() <= 2.0


@pytest.mark.parametrize("e", [-0.62, -0.5, -0.3, 0.25, 1.0])
def test_pow_within_2_ulp(mlib, e):
    rng = np.random.default_rng(2)
    g = 10.0 ** rng.uniform(-3, 5, 300000)
    y = np.empty_like(g)
    mlib.t_pow(_p(g), e, _p(y), len(g))
    assert _ulp_err(y, np.power(g, e)).max() <= 2.0


def test_sincos_within_2_ulp(mlib):
    rng = np.random.default_rng(3)
    R = np.concatenate([rng.uniform(0, 1, 600000), np.arange(0, 1, 1 / 64.0), [0.25, 0.5, 0.75, 0.125]])
    a = 6.283185307179586 * R
    s, c = np.empty_like(a), np.empty_like(a)
    mlib.t_sincos(_p(a), _p(s), _p(c), len(a))
    for got, ref in ((s, np.sin(a)), (c, np.cos(a))):
        big = np.abs(ref) > 1e-3
        assert _ulp_err(got[big], ref[big]).max() <= 2.0
        assert np.abs(got[~big] - ref[~big]).max() <= 1e-16
