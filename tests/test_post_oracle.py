"""Pins the oracle's field quantities of the post-processing modules
(src/postProcessMod.cpp:7-151) with values whose result is exact: moduli on the
axes, the signed zeros of atan2, logarithms of powers of two (log 2^k = k log 2
is exact in binary), and the clamps of LogIntensityUint8 in both directions."""
import numpy as np
import pytest

import pyoracle as O


def test_quantities_closed_forms():
    u = np.array([3 + 4j, -1 + 0j, complex(-1.0, -0.0), 0j, -2j])
    assert np.array_equal(O.field_quantity(u, 0), [5.0, 1.0, 1.0, 0.0, 2.0])
    assert np.array_equal(O.field_quantity(u, 1), [25.0, 1.0, 1.0, 0.0, 4.0])
    arg = O.field_quantity(u, 2)
    assert arg[0] == pytest.approx(np.arctan(4.0 / 3.0), rel=1e-15)
    assert arg[1] == np.pi and arg[2] == -np.pi  # atan2 keeps the sign of a zero imaginary part
    assert arg[3] == 0.0 and arg[4] == -np.pi / 2
    assert np.array_equal(O.field_quantity(u, 3), [3.0, -1.0, -1.0, 0.0, 0.0])


def test_quantity_resizes_matrices_only():
    m = np.asfortranarray(np.arange(60.0).reshape(6, 10, order="F") + 0j)
    r = O.field_quantity(m, 3, 2, 3)
    assert r.shape == (2, 3)
    # block (0, 0): dY = 6/3, dX = 10/4, rows 0..2 and columns 0..2 (inclusive, truncated
    # bounds) of the ramp r + 6 c; block (1, 2): rows 2..4, columns 5..7
    assert r[0, 0] == np.mean([rr + 6 * c for rr in range(3) for c in range(3)])
    assert r[1, 2] == np.mean([rr + 6 * c for rr in range(2, 5) for c in range(5, 8)])
    assert O.field_quantity(m, 3, 6, 3).shape == (6, 10)  # an export size reaching the matrix: kept
    assert O.field_quantity(m[:, 0], 3, 2, 3).shape == (6,)  # vectors are not resized
    assert O.field_quantity(np.zeros((4, 4, 3), complex), 0, 2, 2).shape == (4, 4, 3)


def _on_axes(mod):
    # moduli exactly representable as complex values on the four axes
    axes = np.array([1, 1j, -1, -1j])
    return np.asarray(mod, dtype=float) * axes[np.arange(len(mod)) % 4]


def test_intensity_uint8_is_truncated_ratio():
    u = _on_axes(np.arange(256.0)).reshape(16, 4, 4, order="F")
    assert np.array_equal(O.intensity_uint8(u).ravel(order="F"), np.arange(256))
    # truncation, not rounding: 255 * 1.9 / 2 = 242.25 and 255 * 0.01 / 2 = 1.275
    v = O.intensity_uint8(_on_axes([2.0, 1.9, 0.01, 0.0]))
    assert list(v) == [255, 242, 1, 0]
    # a field of zeros: 0 / 0, which the reference converts with undefined behaviour
    assert list(O.intensity_uint8(np.zeros(3, complex))) == [0, 0, 0]


def test_log_intensity_uint8_powers_of_two():
    u = _on_axes([1.0, 16.0, 256.0, 4.0])
    # bot = log 1 = 0, top = log 256 = 8 log 2: 16 -> 255 / 2 = 127.5 -> 127, 4 -> 63.75 -> 63
    assert list(O.log_intensity_uint8(u)) == [0, 127, 255, 63]
    # setMaxValue(16) lowers top to log 16: 256 clamps to 255, 4 -> 127
    assert list(O.log_intensity_uint8(u, max_value=16.0)) == [0, 255, 255, 127]
    # setMaxValue above the field's maximum leaves top at log 256
    assert list(O.log_intensity_uint8(u, max_value=1024.0)) == [0, 127, 255, 63]
    # setMinValue(4) raises bot to log 4: 1 clamps to 0; 16 -> 255 (2 log 2) / (6 log 2)
    v = O.log_intensity_uint8(u, min_value=4.0)
    assert v[0] == 0 and v[2] == 255 and v[3] == 0 and v[1] in (84, 85)
    # setMinValue below the field's minimum leaves bot at log 1
    assert list(O.log_intensity_uint8(u, min_value=0.25)) == [0, 127, 255, 63]
    # both set: bot = log 1 (field minimum 0.5 raised), top = log 16
    w = _on_axes([0.5, 1.0, 4.0, 16.0, 256.0])
    assert list(O.log_intensity_uint8(w, min_value=1.0, max_value=16.0)) == [0, 0, 127, 255, 255]


def test_log_intensity_uint8_degenerate_fields():
    # uniform modulus: top = bot, every value is 0/0 -> 0
    assert list(O.log_intensity_uint8(_on_axes([2.0, 2.0, 2.0]))) == [0, 0, 0]
    # a zero cell: bot = -inf, 0/0 -> 0 for it, the other cells give inf/inf -> 0 as well
    assert list(O.log_intensity_uint8(_on_axes([0.0, 1.0, 2.0]))) == [0, 0, 0]
    # with setMinValue the zero cell clamps to 0 and the rest map normally
    assert list(O.log_intensity_uint8(_on_axes([0.0, 1.0, 2.0]), min_value=1.0)) == [0, 0, 255]
