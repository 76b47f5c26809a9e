"""CPU checks pinning the far-field restatement in oracle/pyoracle.py
(post::FarField::result, src/farFieldPost.cpp:141-230; resizeMatrix,
src/postProcessing.cpp:29-49) with closed forms: a point scatterer, plane waves
along each axis (fftshift direction, row/column orientation, FFT sign, the
mirrored x axis that Armadillo's conjugating t() gives the row pass), a tilted
wave against the angle-to-index map, block means of simple matrices, and the
form factor of a sphere (Fourier slice of a uniform ball)."""
import numpy as np
import pytest

import pyoracle as O


def full_window_k(dx):
    # k dx = pi puts +-90 degrees exactly at |v| = M/2: the whole spectrum
    return np.pi / dx


def test_point_scatterer_is_flat_one_over_M():
    n, M, dx = 8, 16, 0.5
    sol = np.zeros((n, n), complex)
    sol[3, 5] = 1.0
    ff = O.far_field(sol, None, M, full_window_k(dx), dx, dx, -90.0, 90.0)
    assert ff.shape == (M, M)
    assert np.allclose(ff, 1.0 / M, rtol=1e-14, atol=0)


def test_reference_is_subtracted():
    rng = np.random.default_rng(0)
    n, dx = 16, 0.5
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    ref = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    f1 = O.far_field(a + ref, ref, 32, full_window_k(dx), dx, dx, -90.0, 90.0)
    f2 = O.far_field(a, None, 32, full_window_k(dx), dx, dx, -90.0, 90.0)
    assert np.allclose(f1, f2, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("p", [1, 3, -2])
def test_plane_wave_along_y_lands_on_its_row(p):
    M = n = 16
    dx = 0.5
    j = np.arange(n)[:, None]
    sol = np.exp(2j * np.pi * p * j / M) * np.ones((1, n))
    ff = O.far_field(sol, None, M, full_window_k(dx), dx, dx, -90.0, 90.0)
    r, c = np.unravel_index(np.argmax(ff), ff.shape)
    # forward FFT e^{-2 pi i f j / M} peaks at f = p; fftshift puts f at f + M/2
    assert (r, c) == ((p + M // 2) % M, M // 2)
    assert ff[r, c] == pytest.approx(M ** 3, rel=1e-12)  # |M * M|^2 / M
    ff[r, c] = 0
    assert np.abs(ff).max() < 1e-18 * M ** 3


@pytest.mark.parametrize("q", [1, 4])
def test_plane_wave_along_x_is_mirrored(q):
    M = n = 16
    dx = 0.5
    i = np.arange(n)[None, :]
    sol = np.ones((n, 1)) * np.exp(2j * np.pi * q * i / M)
    ff = O.far_field(sol, None, M, full_window_k(dx), dx, dx, -90.0, 90.0)
    r, c = np.unravel_index(np.argmax(ff), ff.shape)
    # the row pass transforms conj(row): the peak sits at -q, i.e. column M/2 - q
    assert (r, c) == (M // 2, M // 2 - q)


def test_tilted_wave_peaks_at_its_angle_index():
    # a wave exp(i k sin(theta) y) along y scatters at theta: its spectral peak is
    # the index farFieldAngleToIndx gives theta (a whole number of periods here)
    M = n = 64
    dy = 2.0
    k = 2 * np.pi / 0.7
    f0 = 5  # periods across the grid
    sin_t = 2 * np.pi * f0 / (M * dy * k)
    theta = np.degrees(np.arcsin(sin_t))
    y = dy * np.arange(n)[:, None]
    sol = np.exp(1j * k * sin_t * y) * np.ones((1, n))
    ff = O.far_field(sol, None, M, k, dy, dy, -90.0, 90.0)
    r = np.unravel_index(np.argmax(ff), ff.shape)[0]
    ylo = O.far_field_angle_index(-90.0, M, k, dy)
    assert r + ylo == O.far_field_angle_index(theta * (1 + 1e-9), M, k, dy) == f0 + M // 2


def test_angle_index_edges():
    # beyond the resolvable range the index clamps to the ends of the spectrum
    assert O.far_field_angle_index(89.0, 64, 10.0, 1.0) == 63
    assert O.far_field_angle_index(-89.0, 64, 10.0, 1.0) == 0
    assert O.far_field_angle_index(0.0, 64, 10.0, 1.0) == 32
    # truncation toward zero on both sides
    k, dx, M = 1.0, 1.0, 64
    a = np.degrees(np.arcsin(2 * np.pi * 2.5 / (M * k * dx)))
    assert O.far_field_angle_index(a, M, k, dx) == 34
    assert O.far_field_angle_index(-a, M, k, dx) == 30


def test_resize_block_means():
    A = np.full((40, 30), 2.5, order="F")
    assert np.allclose(O.resize_matrix(A, 7, 5), 2.5)
    r = np.arange(40)[:, None] + 1000.0 * np.arange(30)[None, :]
    out = O.resize_matrix(r, 7, 5)
    dY, dX = 40 / 8.0, 30 / 6.0
    for j in range(7):
        for i in range(5):
            r0, r1 = int(j * dY), int((j + 1) * dY)
            c0, c1 = int(i * dX), int((i + 1) * dX)
            # the mean of a ramp over an inclusive block is its midpoint
            assert out[j, i] == pytest.approx((r0 + r1) / 2 + 1000.0 * (c0 + c1) / 2, rel=1e-13)
    # an export size at or above the matrix size keeps it
    assert O.resize_matrix(r, 40, 5).shape == (40, 30)
    assert O.resize_matrix(r, 7, 30).shape == (40, 30)


def test_sphere_form_factor_first_minimum():
    # weak phase object: the difference field is i*k*delta*t(x, y) with t the
    # projected thickness of a ball of radius R; by the Fourier slice theorem its
    # far field is the ball's 3D form factor on the plane qz = 0, whose first zero
    # is at qR = 4.4934.  Along x the pattern is mirrored, a symmetric profile.
    n, dx, R = 256, 1.0, 20.0
    x = dx * (np.arange(n) - n / 2)
    X, Y = np.meshgrid(x, x)
    t = 2 * np.sqrt(np.clip(R ** 2 - X ** 2 - Y ** 2, 0, None))
    sol = 1j * 1e-4 * t
    M = 4096
    k = np.pi / dx
    ff = O.far_field(sol, None, M, k, dx, dx, -90.0, 90.0)
    c = M // 2
    for prof in (ff[c, c:], ff[c:, c], ff[c, c::-1]):
        q = 2 * np.pi * np.arange(prof.size) / (M * dx)
        i_min = np.argmin(prof[: int(6.0 / R / (2 * np.pi / (M * dx)))])
        assert abs(q[i_min] * R - 4.4934) < 0.05
        qr = q[1:40] * R
        fac = (3 * (np.sin(qr) - qr * np.cos(qr)) / qr ** 3) ** 2
        assert np.allclose(prof[1:40] / prof[0], fac, rtol=2e-2, atol=2e-4)


def test_constant_padding_of_a_constant_field_is_a_delta():
    # post::FarField::constantPadding (src/farFieldPost.cpp:314-321) in both passes: a
    # constant field c padded with its own entry 0 is constant over all M entries of
    # every column and then of every row, so the spectrum is c M^2 at the centre and
    # zero elsewhere: |c M^2|^2 / M.  Zero padding of the same field gives a sinc.
    n, M, dx, c = 8, 32, 0.5, 0.7 - 0.2j
    sol = np.full((n, n), c)
    ff = O.far_field(sol, None, M, full_window_k(dx), dx, dx, -90.0, 90.0, padding="constant")
    exp = np.zeros((M, M))
    exp[M // 2, M // 2] = abs(c) ** 2 * M ** 3
    assert np.allclose(ff, exp, rtol=1e-13, atol=1e-9 * exp.max())
    fz = O.far_field(sol, None, M, full_window_k(dx), dx, dx, -90.0, 90.0)
    assert np.count_nonzero(fz > 1e-6 * fz.max()) > 1


def test_constant_padding_repeats_entry_zero():
    # only row 0 non-zero (ones): each column pads with its entry 0 = 1, so its
    # transform is 1 + G(f) with G(f) = sum_{j=n}^{M-1} exp(-2 pi i f j / M); each row of
    # the column pass is then constant (conjugated and padded with itself), so only the
    # centre column is lit, with M |1 + G(f)|^2.  Padding with the last entry (0) would
    # give M everywhere in that column.
    n, M, dx = 8, 32, 0.5
    sol = np.zeros((n, n), complex)
    sol[0, :] = 1.0
    ff = O.far_field(sol, None, M, full_window_k(dx), dx, dx, -90.0, 90.0, padding="constant")
    f = np.arange(M)
    G = np.array([sum(np.exp(-2j * np.pi * ff_ * j / M) for j in range(n, M)) for ff_ in f])
    col = M * np.abs(1.0 + G) ** 2
    col = np.roll(col, M // 2)  # fftshift of an even length
    assert ff[M // 2 + 0, M // 2] == pytest.approx(M * (M - n + 1) ** 2, rel=1e-13)
    assert np.allclose(ff[:, M // 2], col, rtol=1e-12, atol=1e-9 * col.max())
    off = np.delete(ff, M // 2, axis=1)
    assert np.max(off) < 1e-12 * col.max()


@pytest.mark.parametrize("n", [16, 15])
@pytest.mark.parametrize("signed", [False, True])
def test_linear_padding_of_ramp_columns(n, signed):
    # post::FarField::linearPadding (src/farFieldPost.cpp:265-312) on columns p + q j:
    # the fits reproduce the ramp, so the last L = (M - n)/2 entries continue it
    # backwards to p at index M - 1, and the entries after n carry p + q (9 + i - n),
    # because the second fit reads entries 0..9 again against the grid's last ten
    # positions; for odd M - n the middle entry stays 0.  The backward ramp's
    # abscissa is i - M + 1 taken modulo 2^64 in the reference (an int minus an
    # unsigned 64-bit n_elem): 2^64 - j for j = M - 1 - i < 1024 is a double only as
    # 2^64, so those entries are p + q 2^64 and entry M - 1 is p; with the signed
    # difference (PXG_FIX_LINEAR_PAD) they are p + q (i - M + 1).  Identical columns
    # make every row of the column pass constant (its fit is flat), so only the
    # centre column is lit: M |FFT(column)|^2.
    M, dx, xmin = 64, 0.5, -3.0
    p, q = 0.3 + 0.1j, -0.05 + 0.02j
    j = np.arange(n)
    sol = np.tile((p + q * j)[:, None], (1, n))
    L = (M - n) // 2
    col = np.zeros(M, complex)
    col[:n] = p + q * j
    i = np.arange(n, n + L)
    col[n:n + L] = p + q * (9 + i - n)
    i = np.arange(M - L, M)
    col[M - L:] = p + q * ((i - M + 1) if signed else np.where(i < M - 1, 2.0 ** 64, 0.0))
    exp_col = M * np.abs(np.roll(np.fft.fft(col), M // 2)) ** 2
    if n % 2:  # each constant row also keeps one zero: its sum is (M - 1) times the value
        exp_col *= ((M - 1.0) / M) ** 2
    ff = O.far_field(sol, None, M, full_window_k(dx), dx, dx, -90.0, 90.0, padding="linear", xmin=xmin,
                     quirks=O.FIX_LINEAR_PAD if signed else 0)
    assert np.allclose(ff[:, M // 2], exp_col, rtol=1e-9, atol=1e-9 * exp_col.max())
    if n % 2 == 0:  # with the zero, other columns are lit as well
        assert np.max(np.delete(ff, M // 2, axis=1)) < 1e-9 * exp_col.max()


def test_gsl_fit_linear_exact_on_a_line():
    x = O._arma_linspace(-2.0, 7.0, 10)
    c0, c1 = O._gsl_fit_linear(x, 1.5 - 0.25 * x)
    assert c0 == pytest.approx(1.5, abs=1e-14) and c1 == pytest.approx(-0.25, abs=1e-15)
    # least squares of a non-linear set: the normal equations' solution
    y = x ** 2
    A = np.vstack([np.ones_like(x), x]).T
    ref = np.linalg.lstsq(A, y, rcond=None)[0]
    assert np.allclose(O._gsl_fit_linear(x, y), ref, rtol=1e-13)


def test_linear_padding_wrap_staircase():
    """Segments longer than 1024 entries: the wrapped abscissa 2^64 - j is the
    double 2^64 - 2048 r, r = j/2048 rounded half to even (the doubles of
    [2^63, 2^64) are 2048 apart), so a ramp line p + q j pads entry M - 1 - j with
    p + (q/dxf) dxf (2^64 - 2048 r): constant runs, ties at j = 1024, 3072, 5120
    going to r = 0, 2, 2."""
    n, M, dx, xmin = 16, 8192, 0.5, -3.0
    p, q = 0.25 - 0.5j, 0.125 + 0.0625j
    line = (p + q * np.arange(n))[:, None]
    out = O._linear_pad(line, M, 0, xmin, dx)[:, 0]
    L = (M - n) // 2
    dxf = O._arma_linspace(0.0, 10 * dx, 10)[1]
    jj = np.arange(1, L)
    r = np.array([round(j / 2048) for j in jj], dtype=float)  # Python rounds half to even
    for j, want in ((1024, 0), (1025, 1), (3071, 1), (3072, 2), (4000, 2)):
        assert r[j - 1] == want
    exp = p + (q / dxf) * (dxf * (2.0 ** 64 - 2048.0 * r))
    got = out[M - 1 - jj]
    assert np.allclose(got, exp, rtol=1e-15, atol=0)
    assert out[M - 1] == pytest.approx(p, abs=1e-12)
    # the oracle converts (i - M + 1) mod 2^64 directly; the staircase form the device
    # sums in closed form is that conversion exactly, for every j of the segment
    assert all(float((-int(j)) % (1 << 64)) == 2.0 ** 64 - 2048.0 * rj for j, rj in zip(jj, r))
