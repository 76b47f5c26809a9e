"""Pins the CPU FP64 oracle against golden vectors, the reference's own
compiled Thomas solver, the reference's CSG known answers, and closed forms.

Each check is chosen so that a plausible mistake (a dropped term, a sign,
a transposed index) fails it; none retypes the oracle's formula.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
REF_THOMAS = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_thomas.so")


# ----------------------------------------------------------------------- FFT
def test_fft1d_golden_pocketfft():
    g = np.load(os.path.join(GOLD, "fft_golden.npz"))
    for n in (1, 2, 3, 5, 8, 16, 64, 512):
        x = g[f"x1d_{n}"]
        np.testing.assert_allclose(O.fft1d(x, -1), g[f"fwd1d_{n}"], rtol=0, atol=1e-11 * max(n, 1))
        np.testing.assert_allclose(O.fft1d(x, +1), g[f"bwd1d_{n}"], rtol=0, atol=1e-11 * max(n, 1))


def test_fft2d_golden_pocketfft():
    g = np.load(os.path.join(GOLD, "fft_golden.npz"))
    for n in (4, 16, 64):
        a = g[f"x2d_{n}"]
        np.testing.assert_allclose(O.fft2d_colmajor(a, -1), g[f"fwd2d_{n}"], atol=1e-10 * n)
        np.testing.assert_allclose(O.fft2d_colmajor(a, +1), g[f"bwd2d_{n}"], atol=1e-10 * n)


@pytest.mark.parametrize("n", [4, 8, 32, 6, 12])
def test_fft1d_bruteforce_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(j, j) / n)
    np.testing.assert_allclose(O.fft1d(x, -1), W @ x, atol=1e-12 * n)
    np.testing.assert_allclose(O.fft1d(x, +1), W.conj() @ x, atol=1e-12 * n)


def test_fft_parseval_and_impulse():
    n = 256
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X = O.fft1d(x)
    assert abs(np.vdot(X, X).real - n * np.vdot(x, x).real) < 1e-9 * n * n
    e = np.zeros(n, complex)
    e[5] = 1.0
    np.testing.assert_allclose(O.fft1d(e), np.exp(-2j * np.pi * 5 * np.arange(n) / n), atol=1e-13)


# -------------------------------------------------------------------- Thomas
def test_thomas_golden_reference_build():
    """Fixtures produced by the reference's own include/thomasAlgorithm.tpp."""
    g = np.load(os.path.join(GOLD, "thomas_golden.npz"))
    for key in g["cases"]:
        x = O.thomas(g[f"diag_{key}"], g[f"sub_{key}"], g[f"rhs_{key}"])
        np.testing.assert_array_equal(x, g[f"x_{key}"])


@pytest.mark.skipif(not os.path.exists(REF_THOMAS), reason="oracle/_ref not built")
def test_thomas_bitexact_vs_reference_build():
    lib = C.CDLL(REF_THOMAS)
    lib.ref_thomas_c128.argtypes = [C.c_void_p] * 3 + [C.c_uint]
    rng = np.random.default_rng(11)
    for n in (2, 5, 100, 4096):
        d = 3 + rng.standard_normal(n) + 1j * rng.standard_normal(n)
        s = rng.standard_normal(n - 1) + 1j * rng.standard_normal(n - 1)
        r = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        ours = O.thomas(d, s, r)
        dd, ss, rr = d.copy(), s.copy(), r.copy()
        lib.ref_thomas_c128(dd.ctypes.data, ss.ctypes.data, rr.ctypes.data, n)
        np.testing.assert_array_equal(ours, dd)


@pytest.mark.parametrize("n", [2, 3, 17, 64])
def test_thomas_dense_solve(n):
    rng = np.random.default_rng(n + 100)
    d = 4 + rng.standard_normal(n) + 1j * rng.standard_normal(n)
    s = rng.standard_normal(n - 1) + 1j * rng.standard_normal(n - 1)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    A = np.diag(d) + np.diag(s, 1) + np.diag(s, -1)  # symmetric, not Hermitian
    np.testing.assert_allclose(O.thomas(d, s, b), np.linalg.solve(A, b), rtol=1e-12, atol=1e-12)


# ----------------------------------------------------------------------- CSG
def test_csg_transform_kats():
    kat = json.load(open(os.path.join(GOLD, "csg_kat.json")))
    for case in kat["kats"]:
        T = O.mat4_identity()
        op = case["op"]
        if op[0] == "rotate":
            O.rotate(T, op[1], op[2])
        else:
            O.translate(T, *op[1:])
        np.testing.assert_allclose(O.transform(T, *case["p"]), case["expect"], atol=kat["tol"])


def _sphere_scene(r, delta, beta, center=(0.0, 0.0, 0.0)):
    shapes = np.zeros(1, O.SHAPE_DTYPE)
    T = O.mat4_identity()
    O.translate(T, *center)
    shapes[0]["type"], shapes[0]["op"], shapes[0]["T"], shapes[0]["p"] = O.SPHERE, O.UNION, T, (r, 0, 0)
    parts = np.zeros(1, O.PART_DTYPE)
    parts[0]["first_shape"], parts[0]["n_shapes"], parts[0]["op"] = 0, 1, O.UNION
    parts[0]["delta"], parts[0]["beta"] = delta, beta
    modules = np.zeros(1, O.MODULE_DTYPE)
    modules[0]["first_part"], modules[0]["n_parts"] = 0, 1
    return O.CSGMaterial(shapes, parts, modules)


def test_csg_sphere_membership_analytic():
    m = _sphere_scene(4.0, 5e-6, 1e-7, center=(1.0, -2.0, 0.5))
    rng = np.random.default_rng(5)
    for _ in range(500):
        p = rng.uniform(-6, 6, 3)
        inside = np.linalg.norm(p - np.array([1.0, -2.0, 0.5])) < 4.0
        assert m.eval(*p) == ((5e-6, 1e-7) if inside else (0.0, 0.0))


def test_csg_part_and_module_order():
    """Part: last shape decides (difference carves); module: last part wins;
    first module that claims the point wins; reference run -> 0."""
    shapes = np.zeros(3, O.SHAPE_DTYPE)
    for i in range(3):
        shapes[i]["T"] = O.mat4_identity()
    shapes[0]["type"], shapes[0]["p"] = O.BOX, (10, 10, 10)
    shapes[1]["type"], shapes[1]["op"], shapes[1]["p"] = O.SPHERE, O.DIFFERENCE, (2, 0, 0)
    shapes[2]["type"], shapes[2]["p"] = O.SPHERE, (1, 0, 0)
    O.translate(shapes[2]["T"], 3.0, 0.0, 0.0)
    parts = np.zeros(3, O.PART_DTYPE)
    parts[0] = (0, 2, O.UNION, 0, 1e-5, 1e-6)      # box minus sphere r=2
    parts[1] = (2, 1, O.UNION, 0, 7e-6, 7e-7)      # small sphere at x=3 (later part wins)
    parts[2] = (0, 1, O.UNION, 0, 9e-6, 9e-7)      # second module: the whole box
    modules = np.zeros(2, O.MODULE_DTYPE)
    modules[0], modules[1] = (0, 2), (2, 1)
    m = O.CSGMaterial(shapes, parts, modules)
    assert m.eval(0, 0, 0) == (9e-6, 9e-7)          # carved out of module 0 -> module 1 claims
    assert m.eval(4.0, 4.0, 0) == (1e-5, 1e-6)      # box region of module 0
    assert m.eval(3.2, 0, 0) == (7e-6, 7e-7)        # later part overrides
    assert m.eval(20, 0, 0) == (0.0, 0.0)
    ref = O.CSGMaterial(shapes, parts, modules, reference_run=True)
    assert ref.eval(4.0, 4.0, 0) == (0.0, 0.0)


# ----------------------------------------------------------------- sources
def test_gaussian_closed_form_at_waist():
    k = 2 * np.pi / 0.1
    gb = O.Gauss(k, 1.0, 0.0, 3000.0, 0.0, 0.0, 0.0, 0.0)
    g = O.grid(-10000, 10000 - 39.0625, 39.0625, 0, 10, 10, k)
    f = O.gauss3d_field(g, gb, 8)
    xs = -10000 + 39.0625 * np.arange(8)
    ref = np.exp(-(xs[None, :] ** 2 + xs[:, None] ** 2) / 3000.0 ** 2)
    np.testing.assert_allclose(f, ref, rtol=1e-14)
    pw = O.Plane(k, 2.0, 0.0, 0.01)
    np.testing.assert_allclose(O.plane_value(pw, 3.0), 2 * np.exp(1j * k * 3.0 * np.sin(0.01)), rtol=1e-14)


def test_node_count_truncation():
    assert O.nodes(-500, 500, 1000 / 1023) == 1024
    assert O.nodes(-10000, -10000 + 511 * 39.0625, 39.0625) == 512
    # (max - min)/step + 1 truncated to unsigned (src/paraxialSimulation.cpp:107-110): a
    # quotient that rounds just below 511 gives 511 nodes
    assert O.nodes(0, 20000, 20000 / 511) in (511, 512)


# ------------------------------------------------------------ cube storage
def test_cube_block_average_matches_reshape_mean():
    rng = np.random.default_rng(9)
    n, ds = 16, 4
    f = np.asfortranarray(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    cube = np.zeros((n // ds, n // ds, 3), complex, order="F")
    O.cube_store(f, cube, 2, 4)   # step 4 with dsz 2 -> slice 2
    ref = f.reshape(n // ds, ds, n // ds, ds).mean(axis=(1, 3))
    np.testing.assert_allclose(cube[:, :, 2], ref, rtol=1e-14)
    assert np.all(cube[:, :, :2] == 0)


# ------------------------------------------------------------ FFT3D forms
def _fft3d_grid(n, dz, k=2 * np.pi / 0.1, dx=39.0625):
    xmin = -dx * n / 2
    return O.grid(xmin, xmin + (n - 1) * dx, dx, -5000, 5000, dz, k)


def test_fft3d_free_space_composes():
    """delta = beta = 0: n steps of dz equal one step of n*dz (kernel is a semigroup)."""
    n = 32
    rng = np.random.default_rng(1)
    u0 = np.asfortranarray(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    a = O.fft3d_run(_fft3d_grid(n, 10.0), O.NoMaterial, u0, 1, 8)
    b = O.fft3d_run(_fft3d_grid(n, 80.0), O.NoMaterial, u0, 1, 1)
    assert O.rel_l2(a, b) < 1e-13
    # and it is not the identity (the kernel really acts)
    assert O.rel_l2(a, u0) > 1e-3


def _box_material(delta, beta, zlo=-1e9, zhi=1e9):
    shapes = np.zeros(1, O.SHAPE_DTYPE)
    T = O.mat4_identity()
    O.translate(T, 0.0, 0.0, 0.5 * (zlo + zhi))
    shapes[0]["type"], shapes[0]["T"], shapes[0]["p"] = O.BOX, T, (1e12, 1e12, zhi - zlo)
    parts = np.zeros(1, O.PART_DTYPE)
    parts[0] = (0, 1, O.UNION, 0, delta, beta)
    modules = np.zeros(1, O.MODULE_DTYPE)
    modules[0] = (0, 1)
    return O.CSGMaterial(shapes, parts, modules)


def test_fft3d_homogeneous_closed_form():
    n, k, dz, d, b = 16, 2 * np.pi / 0.1, 10.0, 5e-6, 1e-7
    u0 = np.ones((n, n), complex, order="F")
    u = O.fft3d_run(_fft3d_grid(n, dz), _box_material(d, b), u0, 1, 7)
    np.testing.assert_allclose(u, np.exp(-k * (b + 1j * d) * dz * 7), rtol=1e-13)


def test_fft3d_crossing_integral_skips_last_sample():
    """Material switches on at zb inside (z0, z1): only samples i >= 4 of the
    n = 10 sub-samples see it, so the reference's trapezoid (which drops sample
    n-1) gives delta_eff = D*(1 + 2*(n-5))/(2n) = 11/20 D; the full trapezoid
    gives 13/20 D."""
    n, k, dz, D, B = 8, 2 * np.pi / 0.1, 10.0, 4e-6, 2e-8
    g = _fft3d_grid(n, dz)
    z0 = g.zmin  # step 1: z0 = zmin, z1 = zmin + dz
    zb = z0 + 3.5  # between sub-samples 3 (z0+3) and 4 (z0+4)
    mat = _box_material(D, B, zlo=zb, zhi=1e9)
    u0 = np.ones((n, n), complex, order="F")
    u = O.fft3d_run(g, mat, u0, 1, 1)
    eff = 11.0 / 20.0
    np.testing.assert_allclose(u, np.exp(-k * (eff * B + 1j * eff * D) * dz), rtol=1e-13)
    u_fix = O.fft3d_run(g, mat, u0, 1, 1, quirks=O.FIX_REFR_SAMPLES)
    eff = 13.0 / 20.0
    np.testing.assert_allclose(u_fix, np.exp(-k * (eff * B + 1j * eff * D) * dz), rtol=1e-13)


def test_fft3d_absorber_mask():
    """Absorber on a uniform vacuum field equals the separable edge mask with
    factor exp(-(npix-i-1) inv), innermost cell 1, y (columns) then x."""
    n = 16
    g = _fft3d_grid(n, 10.0)
    ab = O.Absorber3D(3, 2, 0.5, 0.25)
    u = O.fft3d_run(g, O.NoMaterial, np.ones((n, n), complex, order="F"), 1, 1, absorber=ab)
    def edge(npix, inv):
        m = np.ones(n)
        for i in range(npix):
            m[i] = m[n - 1 - i] = np.exp(-(npix - i - 1) * inv)
        return m
    mask = edge(2, 0.25)[:, None] * edge(3, 0.5)[None, :]  # rows=y use npix_y
    np.testing.assert_allclose(u, mask, rtol=1e-13)


def test_fft2d_homogeneous_and_free_space():
    k, nx = 2 * np.pi / 0.1, 64
    g = O.grid(-500, 500, 1000 / 63, 0, 1000, 10.0, k)
    mat = O.SlabMaterial(-1e9, 1e9, 3e-6, 2e-7, 0, 0)   # core everywhere
    u = O.fft2d_run(g, mat, np.ones(nx, complex), 1, 5)
    np.testing.assert_allclose(u, np.exp(-k * (2e-7 + 3e-6j) * 10.0 * 5), rtol=1e-13)
    rng = np.random.default_rng(2)
    u0 = rng.standard_normal(nx) + 1j * rng.standard_normal(nx)
    a = O.fft2d_run(g, O.NoMaterial, u0, 1, 4)
    g4 = O.grid(-500, 500, 1000 / 63, 0, 1000, 40.0, k)
    b = O.fft2d_run(g4, O.NoMaterial, u0, 1, 1)
    assert O.rel_l2(a, b) < 1e-13


# ---------------------------------------------------------------- CN forms
def test_cn_dirichlet_vacuum_is_unitary():
    k, nx = 2 * np.pi / 0.1, 200
    g = O.grid(-500, 500, 1000 / (nx - 1), 0, 1e6, 500.0, k)
    rng = np.random.default_rng(4)
    u0 = rng.standard_normal(nx) + 1j * rng.standard_normal(nx)
    u = O.cn_run(g, O.NoMaterial, u0, 1, 50, bc=O.BC_DIRICHLET)
    assert abs(np.linalg.norm(u) / np.linalg.norm(u0) - 1) < 1e-12
    assert O.rel_l2(u, u0) > 1e-2


@pytest.mark.parametrize("bc", [O.BC_TRANSPARENT])
def test_cn_uniform_closed_form(bc):
    """Uniform field + uniform material + TBC: the Laplacian and the boundary
    terms cancel, u <- u (1 - r/2 (b + i d)) / (1 + r/2 (b + i d)) everywhere."""
    k, nx, dz, d, b = 2 * np.pi / 0.1, 64, 500.0, 4.5e-5, 3.5e-6
    g = O.grid(-500, 500, 1000 / (nx - 1), 0, 1e6, dz, k)
    mat = O.SlabMaterial(-1e9, 1e9, d, b, 0, 0)
    u = O.cn_run(g, mat, np.ones(nx, complex), 1, 3, bc=bc)
    r = k * dz
    fac = (1 - 0.5 * r * (b + 1j * d)) / (1 + 0.5 * r * (b + 1j * d))
    np.testing.assert_allclose(u, fac ** 3, rtol=1e-12)


def test_cn_absorber_extension():
    k, nx = 2 * np.pi / 0.1, 64
    g = O.grid(-500, 500, 1000 / (nx - 1), 0, 1e6, 500.0, k)
    u = O.cn_run(g, O.NoMaterial, np.ones(nx, complex), 1, 1, bc=O.BC_TRANSPARENT,
                 absorb_npix=5, absorb_inv=0.1)
    m = np.ones(nx)
    for i in range(5):
        m[i] = m[nx - 1 - i] = np.exp(-(5 - i - 1) * 0.1)
    np.testing.assert_allclose(u, m, rtol=1e-12)


# --------------------------------------------------------------- ADI forms
def test_adi_uniform_half_step_closed_forms():
    """Uniform field, uniform material, TBC on: x-pass v = u/(1 + h k (b + i d));
    y-pass (reference parenthesis) w = v (1/h)/(1/h + i/(k dy^2 + k(b+id)) - i/(k dy^2));
    with the physics fix the y-pass equals the x-pass form."""
    n, k, dz, d, b = 16, 2 * np.pi / 0.1, 20.0, 5e-6, 1e-7
    dx = 19.53125
    g = O.grid(-n * dx / 2, -n * dx / 2 + (n - 1) * dx, dx, -5000, 5000, dz, k)
    mat = _box_material(d, b)
    h = dz / 2
    u = O.adi_run(g, mat, np.ones((n, n), complex, order="F"), 1, 1)
    v = 1 / (1 + h * k * (b + 1j * d))
    w = v * (1 / h) / (1 / h + 1j / (k * dx * dx + k * (b + 1j * d)) - 1j / (k * dx * dx))
    np.testing.assert_allclose(u, w, rtol=1e-12)
    u_fix = O.adi_run(g, mat, np.ones((n, n), complex, order="F"), 1, 1, quirks=O.FIX_ADI_Y_DIAG)
    np.testing.assert_allclose(u_fix, v * v, rtol=1e-12)


def test_adi_vacuum_dirichlet_separable_diffusion():
    """delta = beta = 0: the x-pass acts only along x, so a field that is a
    product f(y) g(x) stays separable after one full step."""
    n, k, dz = 16, 2 * np.pi / 0.1, 20.0
    dx = 19.53125
    g = O.grid(0, (n - 1) * dx, dx, 0, 100, dz, k)
    rng = np.random.default_rng(8)
    fy = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    gx = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    u0 = np.asfortranarray(np.outer(fy, gx))
    u = O.adi_run(g, O.NoMaterial, u0, 1, 1, use_tbc=False)
    s = np.linalg.svd(u, compute_uv=False)
    assert s[1] / s[0] < 1e-12


# --------------------------------------------------------- curved waveguide
def test_curved_waveguide_membership():
    w = O.CurvedMaterial(R=1e7, width=100.0, clad_delta=4.5e-5, clad_beta=3.5e-6, dx=1.0, dz=100.0)
    f = C.cast(O.lib().orc_curved_matfn, C.CFUNCTYPE(None, C.c_void_p, C.c_double, C.c_double,
                                                     C.c_double, C.c_void_p, C.c_void_p))
    def ev(x, z):
        d, b = C.c_double(), C.c_double()
        f(C.cast(C.byref(w.wg), C.c_void_p), x, 0.0, z, C.byref(d), C.byref(b))
        return d.value, b.value
    # deep inside at z = 0: core (inside material defaults to 0)
    assert ev(50.0, 0.0) == (0.0, 0.0)
    # near the inner border: averaged
    assert ev(0.5, 0.0) == (0.5 * 4.5e-5, 0.5 * 3.5e-6)
    # outside
    assert ev(-10.0, 0.0) == (4.5e-5, 3.5e-6)
    # the guide bends: at z the inner edge sits at x = -z^2/(2R)
    z = 1e5
    x_edge = -z * z / (2 * 1e7)
    assert ev(x_edge + 50.0, z) == (0.0, 0.0)
    assert ev(x_edge - 5.0, z) == (4.5e-5, 3.5e-6)
