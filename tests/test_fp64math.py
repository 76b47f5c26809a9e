"""CPU checks of the FP64 exp and sincos that the FFT passes inline for their
phase screens (paxpro_amd/csrc/pxg_fp64math.hpp).  The header is compiled for
the host with g++ (every multiply feeding an add is an explicit fma, so the
host build rounds as the device build does) and compared with the C library
(numpy's exp/sin/cos), which is correctly rounded or within an ulp."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = r"""
#include "pxg_fp64math.hpp"
extern "C" void exp_many(const double *x, double *y, long n)
{
  for (long i = 0; i < n; ++i) y[i] = pxg::exp_fp64(x[i]);
}
extern "C" void sincos_many(const double *x, double *s, double *c, long n)
{
  for (long i = 0; i < n; ++i) pxg::sincos_fp64(x[i], s[i], c[i]);
}
"""


@pytest.fixture(scope="module")
def mathlib(tmp_path_factory):
    d = tmp_path_factory.mktemp("fp64math")
    src, so = d / "m.cpp", d / "libm64.so"
    src.write_text(SRC)
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I",
                    os.path.join(ROOT, "paxpro_amd", "csrc"), str(src), "-o", str(so)], check=True)
    L = C.CDLL(str(so))
    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    L.exp_many.argtypes = [dp, dp, C.c_long]
    L.sincos_many.argtypes = [dp, dp, dp, C.c_long]
    return L


def exp_of(L, x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    L.exp_many(x, y, x.size)
    return y


def sincos_of(L, x):
    x = np.ascontiguousarray(x, np.float64)
    s, c = np.empty_like(x), np.empty_like(x)
    L.sincos_many(x, s, c, x.size)
    return s, c


def test_exp_within_an_ulp_over_the_normal_range(mathlib):
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-708.0, 709.0, 200_000), rng.uniform(-1.0, 1.0, 50_000),
                        np.array([0.0, -0.0, 1e-300, -1e-300, 0.34657359027997264,
                                  -0.34657359027997264, np.log(2.0), 1.0, -1.0])])
    y, ref = exp_of(mathlib, x), np.exp(x)
    rel = np.abs(y - ref) / ref
    assert rel.max() <= 2.3e-16, rel.max()
    assert exp_of(mathlib, [0.0])[0] == 1.0


def test_exp_edges(mathlib):
    y = exp_of(mathlib, [-745.0, -720.0, -746.5, -1e6, 710.5, 1e6, np.nan, -np.inf])
    ref = np.exp(np.array([-745.0, -720.0]))
    assert np.all(np.abs(y[:2] - ref) <= 1e-323)  # subnormal results: absolute
    assert y[2] == 0.0 and y[3] == 0.0 and y[7] == 0.0
    assert np.isinf(y[4]) and np.isinf(y[5])
    assert np.isnan(y[6])


@pytest.mark.parametrize("mag", [1.0, np.pi, 1e3, 1e6, 1e9, 1e12, 1e15])
def test_sincos_absolute_error(mathlib, mag):
    rng = np.random.default_rng(int(np.log10(mag) * 7) + 3)
    x = rng.uniform(-mag, mag, 100_000)
    s, c = sincos_of(mathlib, x)
    assert np.abs(s - np.sin(x)).max() <= 2.3e-16
    assert np.abs(c - np.cos(x)).max() <= 2.3e-16


def test_sincos_quadrants_and_special_points(mathlib):
    # every quadrant, the reduction boundaries and the multiples of pi/2
    q = np.arange(-4000, 4001, dtype=np.float64)
    h = np.pi / 2
    x = np.concatenate([q * h, q * h + np.pi / 4, q * h - np.pi / 4, q * h + 1e-9,
                        np.array([0.0, -0.0, 1e-300, 0.5, -0.5, 2.0, -2.0])])
    s, c = sincos_of(mathlib, x)
    assert np.abs(s - np.sin(x)).max() <= 2.3e-16
    assert np.abs(c - np.cos(x)).max() <= 2.3e-16
    # closed forms: sin^2 + cos^2 = 1, sin(pi/6) = 1/2, cos(pi/3) = 1/2
    assert np.abs(s * s + c * c - 1.0).max() <= 4.5e-16
    s6, c3 = sincos_of(mathlib, [np.pi / 6])[0][0], sincos_of(mathlib, [np.pi / 3])[1][0]
    assert abs(s6 - 0.5) <= 1.2e-16 and abs(c3 - 0.5) <= 1.2e-16
    s, c = sincos_of(mathlib, [np.inf, -np.inf, np.nan])
    assert np.all(np.isnan(s)) and np.all(np.isnan(c))


def test_screen_factor_matches_complex_exp(mathlib):
    # the screen exp(-k (beta + i delta) dz) as the passes form it, against numpy's
    # complex exponential, over the span of X-ray and optical materials
    rng = np.random.default_rng(5)
    k = 2 * np.pi / rng.uniform(0.01, 1.0, 20_000)
    dz = rng.uniform(0.1, 100.0, k.size)
    delta = 10.0 ** rng.uniform(-8, -2, k.size)
    beta = 10.0 ** rng.uniform(-10, -3, k.size)
    amp = exp_of(mathlib, (-k * beta) * dz)
    s, c = sincos_of(mathlib, (-k * delta) * dz)
    ours = amp * (c + 1j * s)
    ref = np.exp(-k * (beta + 1j * delta) * dz)
    assert np.abs(ours - ref).max() <= 1e-15 * np.abs(ref).max() + 1e-300
