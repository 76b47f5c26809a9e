"""Closed forms that fix the diffraction and diffusion coefficients of the
oracle's four propagators, with dy != dx wherever the solver has two axes.

The older oracle checks (free-space composition, Cayley unitarity, uniform
fields, separability) hold for any real multiple of the Laplacian, so a wrong
coefficient, sign or axis would pass them.  These do not:

* FFT3D / FFT2D: a grid plane wave exp(2 pi i (p c + q r)/N) is an eigenvector
  of the split step in vacuum with eigenvalue exp(-i dz (kx^2 + ky^2)/(2k)),
  kx = 2 pi min(p, N-p)/(dx N) from the column index, ky from the row index and
  dy (src/fftSolver3D.cpp:24-50,103-115; src/fftSolver2D.cpp:15-19,44-51).
* CN (Dirichlet, vacuum): sin(pi m (j+1)/(N+1)) is multiplied each step by
  (1 - (i rho/2)(1 - cos t))/(1 + (i rho/2)(1 - cos t)), t = pi m/(N+1),
  rho = dz/(k dx^2) (src/crancNicholson.cpp:25-54,83-88).
* ADI without the transparent boundary: a product of sine modes is an
  eigenvector of both half-steps, in vacuum and in a homogeneous medium, with
  the per-half-step factors written out below for both settings of the y-pass
  parenthesis (src/alternatingDirectionSolver.cpp:38-51,103-116).
* Physics: a Gaussian beam in free space follows the paraxial closed form
  u = (1 + i z/zR)^(-1/2) exp(-x^2/(w0^2 (1 + i z/zR))) per transverse axis,
  zR = k w0^2/2, so its width is w0 sqrt(1 + (z/zR)^2).

Each of these mutations of oracle/paxoracle.c was run against this file and
fails at least one test here (they all passed the older checks):
  FFT3D kernel coefficient doubled (0.5*dz/k -> 1.0*dz/k); FFT3D kernel sign
  flipped; kx taken from the row index and ky from the column index; FFT2D
  kernel coefficient doubled; CN rho doubled; ADI x-implicit diffusion doubled
  consistently, transparent-boundary terms included.
"""
import numpy as np
import pytest

import pyoracle as O

K = 2 * np.pi / 0.1


def _grid2(n, dx, dy, dz, nsteps):
    xmin, ymin = -dx * n / 2, -dy * n / 2
    return O.grid(xmin, xmin + (n - 1) * dx, dx, 0.0, nsteps * dz, dz, K,
                  ymin=ymin, ymax=ymin + (n - 1) * dy, dy=dy)


def _kfold(i, n, d):
    m = min(i, n - i)
    return 2 * np.pi * m / (d * n)


def _box(delta, beta):
    shapes = np.zeros(1, O.SHAPE_DTYPE)
    shapes[0]["type"], shapes[0]["T"], shapes[0]["p"] = O.BOX, O.mat4_identity(), (1e12, 1e12, 1e12)
    parts = np.zeros(1, O.PART_DTYPE)
    parts[0] = (0, 1, O.UNION, 0, delta, beta)
    modules = np.zeros(1, O.MODULE_DTYPE)
    modules[0] = (0, 1)
    return O.CSGMaterial(shapes, parts, modules)


# ------------------------------------------------------------- split step
@pytest.mark.parametrize("p,q", [(1, 0), (0, 3), (5, 2), (11, 13), (8, 8), (15, 1)])
def test_fft3d_plane_wave_eigenmode_anisotropic(p, q):
    """Vacuum, dx != dy: u0 = exp(2 pi i (p c + q r)/N) comes back multiplied by
    exp(-i s dz (kx^2 + ky^2)/(2k)) after s steps, kx from (p, dx), ky from (q, dy)."""
    n, dx, dy, dz, s = 16, 39.0625, 23.0, 2000.0, 3
    g = _grid2(n, dx, dy, dz, s)
    r = np.arange(n)[:, None]
    c = np.arange(n)[None, :]
    u0 = np.asfortranarray(np.exp(2j * np.pi * (p * c + q * r) / n))
    u = O.fft3d_run(g, O.NoMaterial, u0, 1, s)
    kx, ky = _kfold(p, n, dx), _kfold(q, n, dy)
    lam = np.exp(-1j * s * dz * (kx * kx + ky * ky) / (2 * K))
    np.testing.assert_allclose(u, lam * u0, rtol=0, atol=1e-12)
    if (p, q) == (5, 2):
        # the phase is far from a multiple of 2 pi, so the check has teeth
        assert abs(lam - 1) > 0.1


@pytest.mark.parametrize("p", [1, 7, 32, 50, 63])
def test_fft2d_plane_wave_eigenmode(p):
    n, dx, dz, s = 64, 1000 / 63, 3000.0, 4
    g = O.grid(-500, 500, dx, 0.0, s * dz, dz, K)
    j = np.arange(n)
    u0 = np.exp(2j * np.pi * p * j / n)
    u = O.fft2d_run(g, O.NoMaterial, u0, 1, s)
    kx = _kfold(p, n, dx)
    lam = np.exp(-1j * s * dz * kx * kx / (2 * K))
    np.testing.assert_allclose(u, lam * u0, rtol=0, atol=1e-12)
    if p == 7:
        assert abs(lam - 1) > 0.1


# ------------------------------------------------------------- Crank-Nicolson
@pytest.mark.parametrize("m", [1, 4, 17, 40])
def test_cn_dirichlet_sine_modes(m):
    """Vacuum, Dirichlet: the CN matrices share the sine eigenvectors of the
    second difference; each step multiplies mode m by the Cayley factor."""
    n, dz, s = 47, 500.0, 5
    dx = 1000 / (n - 1)
    g = O.grid(-500, 500, dx, 0, 1e6, dz, K)
    j = np.arange(n)
    t = np.pi * m / (n + 1)
    u0 = np.sin(t * (j + 1)).astype(complex)
    u = O.cn_run(g, O.NoMaterial, u0, 1, s, bc=O.BC_DIRICHLET)
    rho = dz / (K * dx * dx)
    a = 0.5j * rho * (1 - np.cos(t))
    gain = (1 - a) / (1 + a)
    np.testing.assert_allclose(u, gain ** s * u0, rtol=0, atol=1e-12)
    assert abs(gain ** s - 1) > 1e-4   # eight orders above the tolerance


# ------------------------------------------------------------- ADI
def _adi_factors(n, mx, my, k, dx, dy, h, delta, beta, fix):
    tx, ty = np.pi * mx / (n + 1), np.pi * my / (n + 1)
    ax = 1j / (k * dx * dx)           # i/(k dx^2)
    ay = 1j / (k * dy * dy)
    m = k * (beta + 1j * delta)
    # x-implicit: diag 1/h + i/(k dx^2) + k(b + i d), off -i/(2k dx^2);
    # rhs (1/h - i/(k dy^2)) u + i/(2k dy^2)(u_{y-1} + u_{y+1})
    fx = (1 / h - ay * (1 - np.cos(ty))) / (1 / h + ax * (1 - np.cos(tx)) + m)
    # y-implicit: rhs (1/h - i/(k dx^2)) u + i/(2k dx^2)(u_{x-1} + u_{x+1})
    num = 1 / h - ax * (1 - np.cos(tx))
    if fix:
        den = 1 / h + ay * (1 - np.cos(ty)) + m
    else:   # diag 1/h + i/(k dy^2 + k b + i k d), off -i/(2k dy^2)
        den = 1 / h + 1j / (k * dy * dy + m) - ay * np.cos(ty)
    return fx, num / den


@pytest.mark.parametrize("mx,my", [(1, 1), (3, 7), (12, 2), (9, 14)])
@pytest.mark.parametrize("medium", [(0.0, 0.0), (3e-4, 2e-5)])
@pytest.mark.parametrize("fix", [0, 1])
def test_adi_sine_mode_products(mx, my, medium, fix):
    """No transparent boundary, dx != dy: u0 = sin(my t_y (r+1)) sin(mx t_x (c+1))
    is an eigenvector of both half-steps, in vacuum and in a homogeneous medium;
    the y-pass factor differs between the reference's parenthesis and its fix
    only where the medium is not vacuum."""
    n, dx, dy, dz = 15, 1.0, 1.5, 100.0
    delta, beta = medium
    g = _grid2(n, dx, dy, dz, 2)
    r = np.arange(n)[:, None]
    c = np.arange(n)[None, :]
    u0 = np.asfortranarray((np.sin(np.pi * my * (r + 1) / (n + 1)) *
                            np.sin(np.pi * mx * (c + 1) / (n + 1))).astype(complex))
    mat = O.NoMaterial if delta == 0 else _box(delta, beta)
    quirks = O.FIX_ADI_Y_DIAG if fix else 0
    u = O.adi_run(g, mat, u0, 1, 1, use_tbc=False, quirks=quirks)
    fx, fy = _adi_factors(n, mx, my, K, dx, dy, dz / 2, delta, beta, fix)
    np.testing.assert_allclose(u, fx * fy * u0, rtol=0, atol=1e-12)
    assert abs(fx * fy - 1) > 1e-2


# ------------------------------------------------------------- Gaussian beams
def _beam_1d(x, w0, z):
    zr = K * w0 * w0 / 2
    q = 1 + 1j * z / zr
    return q ** -0.5 * np.exp(-x * x / (w0 * w0 * q))


def _width(x, inten):
    """2 sqrt(<x^2>) of an intensity profile: the 1/e^2 radius of a Gaussian."""
    return 2 * np.sqrt(np.sum(x * x * inten) / np.sum(inten))


def test_fft3d_astigmatic_gaussian_beam_free_space():
    """Split step in vacuum against the exact paraxial beam: elliptic waist
    (150 nm along x, 250 nm along y) on a grid with dx = 20, dy = 25 nm,
    propagated 1.2 mm (1.7 and 0.6 Rayleigh lengths)."""
    n, dx, dy, dz, s = 256, 20.0, 25.0, 40000.0, 30
    wx, wy = 150.0, 250.0
    g = _grid2(n, dx, dy, dz, s)
    x = g.xmin + dx * np.arange(n)
    y = g.ymin + dy * np.arange(n)
    u0 = np.asfortranarray(np.outer(_beam_1d(y, wy, 0.0), _beam_1d(x, wx, 0.0)))
    u = O.fft3d_run(g, O.NoMaterial, u0, 1, s)
    z = s * dz
    exact = np.outer(_beam_1d(y, wy, z), _beam_1d(x, wx, z))
    assert O.rel_l2(u, exact) < 1e-9
    inten = np.abs(u) ** 2
    zrx, zry = K * wx * wx / 2, K * wy * wy / 2
    assert abs(_width(x, inten.sum(axis=0)) / (wx * np.sqrt(1 + (z / zrx) ** 2)) - 1) < 1e-9
    assert abs(_width(y, inten.sum(axis=1)) / (wy * np.sqrt(1 + (z / zry) ** 2)) - 1) < 1e-9


def test_fft2d_gaussian_beam_free_space():
    n, dx, dz, s = 512, 10.0, 20000.0, 40
    w0 = 120.0
    xmin = -dx * n / 2
    g = O.grid(xmin, xmin + (n - 1) * dx, dx, 0.0, s * dz, dz, K)
    x = xmin + dx * np.arange(n)
    u = O.fft2d_run(g, O.NoMaterial, _beam_1d(x, w0, 0.0), 1, s)
    z = s * dz
    assert O.rel_l2(u, _beam_1d(x, w0, z)) < 1e-9


def test_cn_gaussian_beam_width():
    """Finite differences in x and z: the width matches w0 sqrt(1 + (z/zR)^2)
    to the scheme's O(dx^2, dz^2) error (well under 1 %); a doubled or
    sign-flipped diffusion term is off by tens of percent."""
    w0, dx, dz, s = 100.0, 5.0, 1000.0, 400
    n = 801
    xmin = -dx * (n - 1) / 2
    g = O.grid(xmin, -xmin, dx, 0.0, s * dz, dz, K)
    x = xmin + dx * np.arange(n)
    u = O.cn_run(g, O.NoMaterial, _beam_1d(x, w0, 0.0), 1, s, bc=O.BC_TRANSPARENT)
    z = s * dz
    zr = K * w0 * w0 / 2
    assert abs(_width(x, np.abs(u) ** 2) / (w0 * np.sqrt(1 + (z / zr) ** 2)) - 1) < 2e-3
    assert O.rel_l2(u, _beam_1d(x, w0, z)) < 1e-2


def test_adi_gaussian_beam_width_anisotropic():
    wx, wy, dx, dy, dz, s = 100.0, 140.0, 6.0, 8.0, 2000.0, 100
    n = 160
    g = _grid2(n, dx, dy, dz, s)
    x = g.xmin + dx * np.arange(n)
    y = g.ymin + dy * np.arange(n)
    u0 = np.asfortranarray(np.outer(_beam_1d(y, wy, 0.0), _beam_1d(x, wx, 0.0)))
    u = O.adi_run(g, O.NoMaterial, u0, 1, s)
    z = s * dz
    inten = np.abs(u) ** 2
    zrx, zry = K * wx * wx / 2, K * wy * wy / 2
    assert abs(_width(x, inten.sum(axis=0)) / (wx * np.sqrt(1 + (z / zrx) ** 2)) - 1) < 5e-3
    assert abs(_width(y, inten.sum(axis=1)) / (wy * np.sqrt(1 + (z / zry) ** 2)) - 1) < 5e-3
    exact = np.outer(_beam_1d(y, wy, z), _beam_1d(x, wx, z))
    assert O.rel_l2(u, exact) < 2e-2


# ------------------------------------------------------------- ParaxialEquation
DEFAULT_EQ = [[1, 0, 0, 0], [1, 0, 0, 0], [1, 0, 0, 0], [0, 0, 0, 0]]


def _cn_grid(n, dz):
    dx = 1000 / (n - 1)
    return O.grid(-500, 500, dx, 0, 1e6, dz, K), dx


def test_cn_equation_default_coefficients_are_the_default_path():
    """F = G = H = 1, J = 0 given explicitly reproduces the default equation bit
    for bit (slab material, transparent boundary)."""
    n = 64
    g, _ = _cn_grid(n, 500.0)
    mat = O.SlabMaterial(-100, 100, 0.0, 0.0, 4.5e-5, 3.5e-6)
    u0 = np.exp(-np.linspace(-3, 3, n) ** 2).astype(complex)
    a = O.cn_run(g, mat, u0, 1, 7)
    b = O.cn_run(g, mat, u0, 1, 7, equation=O.PolyEq(DEFAULT_EQ))
    np.testing.assert_array_equal(a, b)


def test_cn_equation_j_shifts_delta():
    """J enters as delta - J (src/crancNicholson.cpp:45-46): a constant J in a
    medium of decrement d is the default equation in a medium of d - J, bit for bit."""
    n, d, b, j = 48, 4.5e-5, 3.5e-6, 1.75e-5
    g, _ = _cn_grid(n, 500.0)
    u0 = np.exp(-np.linspace(-2, 2, n) ** 2).astype(complex)
    eq = O.PolyEq([[1, 0, 0, 0], [1, 0, 0, 0], [1, 0, 0, 0], [j, 0, 0, 0]])
    a = O.cn_run(g, O.SlabMaterial(-1e9, 1e9, d, b, 0, 0), u0, 1, 5, equation=eq)
    ref = O.cn_run(g, O.SlabMaterial(-1e9, 1e9, d - j, b, 0, 0), u0, 1, 5)
    np.testing.assert_array_equal(a, ref)
    assert O.rel_l2(a, O.cn_run(g, O.SlabMaterial(-1e9, 1e9, d, b, 0, 0), u0, 1, 5)) > 1e-3


@pytest.mark.parametrize("m", [1, 5, 20])
def test_cn_equation_constant_coefficients_sine_modes(m):
    """Constant F = c, G = g, H = h, vacuum, Dirichlet: the sine modes stay
    eigenvectors and each step multiplies mode m by (c - a)/(c + a),
    a = (i rho g h/2)(1 - cos t): F scales the z derivative, G H the Laplacian."""
    n, dz, s = 47, 500.0, 4
    c, gg, h = 1.7, 0.6, 1.9
    g, dx = _cn_grid(n, dz)
    t = np.pi * m / (n + 1)
    u0 = np.sin(t * (np.arange(n) + 1)).astype(complex)
    eq = O.PolyEq([[c, 0, 0, 0], [gg, 0, 0, 0], [h, 0, 0, 0], [0, 0, 0, 0]])
    u = O.cn_run(g, O.NoMaterial, u0, 1, s, bc=O.BC_DIRICHLET, equation=eq)
    rho = dz / (K * dx * dx)
    a = 0.5j * rho * gg * h * (1 - np.cos(t))
    np.testing.assert_allclose(u, ((c - a) / (c + a)) ** s * u0, rtol=0, atol=1e-12)
    # and not the default equation's factor
    a0 = 0.5j * rho * (1 - np.cos(t))
    assert abs(((c - a) / (c + a)) ** s - ((1 - a0) / (1 + a0)) ** s) > 1e-5   # seven orders above the tolerance


@pytest.mark.parametrize("m", [2, 9])
def test_cn_equation_z_dependent_coefficients_sine_modes(m):
    """F and G varying with z, H = h, vacuum, Dirichlet: step s (z = s dz, primes at
    z - dz) multiplies mode m by
      (F' + (i rho h/2)(G' cos t - G)) / (F + (i rho h/2) G (1 - cos t)),
    with G, not G', in the centre term of the right-hand side, as the reference
    writes it (src/crancNicholson.cpp:87)."""
    n, dz, s = 31, 500.0, 5
    f0, fz, g0, gz, h = 1.3, 2e-4, 0.8, -1.5e-4, 1.2
    g, dx = _cn_grid(n, dz)
    t = np.pi * m / (n + 1)
    u0 = np.sin(t * (np.arange(n) + 1)).astype(complex)
    eq = O.PolyEq([[f0, 0, fz, 0], [g0, 0, gz, 0], [h, 0, 0, 0], [0, 0, 0, 0]])
    u = O.cn_run(g, O.NoMaterial, u0, 1, s, bc=O.BC_DIRICHLET, equation=eq)
    rho = dz / (K * dx * dx)
    gain = 1.0 + 0j
    for st in range(1, s + 1):
        z = st * dz
        F, G = f0 + fz * z, g0 + gz * z
        Fp, Gp = f0 + fz * (z - dz), g0 + gz * (z - dz)
        gain *= (Fp + 0.5j * rho * h * (Gp * np.cos(t) - G)) / (F + 0.5j * rho * h * G * (1 - np.cos(t)))
    np.testing.assert_allclose(u, gain * u0, rtol=0, atol=1e-12)
