"""CPU checks pinning the transmittivity restatement in oracle/pyoracle.py
(post::Transmittivity, src/transmittivity.cpp:30-83) with closed forms: the
trapezoid rule is exact for intensities linear in x, a uniform field transmits the
ratio of the guide's index spans, and the guide edges follow -z^2/(2R)."""
import numpy as np
import pytest

import pyoracle as O


def test_trapezoid_exact_for_linear_intensity():
    xmin, dx = -5.0, 0.25
    x = xmin + dx * np.arange(80)
    a, b = 2.0, 0.3
    col = np.sqrt(a + b * x) * np.exp(1j * 0.7 * x)  # |u|^2 = a + b x
    s, e = 7, 61
    xs, xe = x[s], x[e]
    exact = (xe - xs) * (a + b * (xs + xe) / 2)
    assert O.trapezoid_intensity(col, s, e, xmin, dx) == pytest.approx(exact, rel=1e-13)
    # degenerate ranges: one cell (0/0), reversed ends (unsigned span)
    assert np.isnan(O.trapezoid_intensity(col, 5, 5, xmin, dx))
    ends = abs(col[9]) ** 2 + abs(col[4]) ** 2
    wrapped = ends * (-5 * dx) / float((1 << 32) - 5) * 0.5
    assert O.trapezoid_intensity(col, 9, 4, xmin, dx) == pytest.approx(wrapped, rel=1e-13)


def test_closest_index_truncates():
    assert O.closest_index(0.99, 0.0, 0.5) == 1
    assert O.closest_index(1.0, 0.0, 0.5) == 2
    assert O.closest_index(-0.74, -1.0, 0.5) == 0
    with pytest.raises(ValueError):
        O.closest_index(-1.01, -1.0, 0.5)


def test_uniform_field_transmits_span_ratio():
    xmin, dx, nx = -400.0, 0.5, 1000
    zmin, dz, nz = 0.0, 20.0, 11
    R, width = 4e5, 50.0
    sol = np.ones((nx, nz), complex)
    T, i0 = O.transmittivity_curved(sol, xmin, dx, zmin, dz, R, width)
    s0, e0 = O.closest_index(0.0, xmin, dx), O.closest_index(width, xmin, dx)
    assert i0 == pytest.approx((e0 - s0) * dx, rel=1e-14)
    for n in range(1, nz):
        z = zmin + n * dz
        xs = -0.5 * z * z / R
        s, e = O.closest_index(xs, xmin, dx), O.closest_index(xs + width, xmin, dx)
        assert T[n - 1] == pytest.approx((e - s) / (e0 - s0), rel=1e-14)


def test_guide_follows_the_bend():
    # intensity 1 inside the bent guide at each step, 0 outside: the sum picks the
    # guide's cells at each z and the transmittivity stays near 1 as the guide moves
    xmin, dx, nx = -600.0, 1.0, 800
    zmin, dz, nz = 0.0, 100.0, 9
    R, width = 2e4, 40.0
    x = xmin + dx * np.arange(nx)
    sol = np.zeros((nx, nz), complex)
    for n in range(nz):
        z = zmin + n * dz
        xs = -0.5 * z * z / R
        sol[(x >= xs - dx) & (x <= xs + width + dx), n] = 1.0
    sol[:, 1] = ((x >= -dx) & (x <= width + dx)).astype(float)  # step 1 sits at z = 0 extent
    T, _ = O.transmittivity_curved(sol, xmin, dx, zmin, dz, R, width)
    assert np.all(np.abs(T - 1.0) < 0.06)
