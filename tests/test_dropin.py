"""The C++ drop-in classes (paxpro_amd/cpp) driven like the reference's
examples, compared with the CPU oracle on the same inputs.  The driver binary
is built by __graft_entry__.build() (tests/dropin/Makefile)."""
import os
import pathlib
import subprocess

import numpy as np
import pytest

import pyoracle as O
from gpu_common import announce, rel_l2
from scenes import K_01, gaussian_field, spheres_scene

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "dropin", "dropin_driver")


def run(tmp_path, *args):
    out = str(tmp_path / "out.bin")
    cmd = [DRIVER, args[0], str(args[1]), str(args[2]), out] + [str(a) for a in args[3:]]
    announce("dropin", " ".join(cmd[1:4]))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return out


def load(path, shape, order="F"):
    return np.fromfile(path, dtype=np.complex128).reshape(shape, order=order)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="drop-in driver not built")
def test_dropin_transform_kats_cpu():
    r = subprocess.run([DRIVER, "kat"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "KAT OK" in r.stdout


def _grid3d(n, nsteps):
    extent, dz = 20000.0, 10.0
    dx = extent / n
    xmin = -extent / 2
    zmin = -nsteps * dz / 2
    return O.grid(xmin, xmin + (n - 1) * dx, dx, zmin, zmin + nsteps * dz, dz, K_01)


@pytest.mark.gpu
@pytest.mark.parametrize("mode,n", [("fft3d", 64), ("fft3d_user", 64), ("fft3d_step", 64), ("fft3d_abs", 64),
                                    ("fft3d_reset", 64), ("fft3d", 100), ("fft3d_user", 100)])
def test_dropin_fft3d(tmp_path, mode, n):
    nsteps = 8
    out = run(tmp_path, mode, n, nsteps)
    g = _grid3d(n, nsteps)
    u0 = gaussian_field(g, n, 3000.0)
    csg = O.CSGMaterial(*spheres_scene([(0.0, 0.0, 0.0, 4000.0, 5e-6, 1e-7)]))
    ab = None
    if mode == "fft3d_abs":
        npix = int(600.0 / g.dx + 1)
        ab = O.Absorber3D(npix, npix, g.dx / 300.0, g.dx / 300.0)
    dsz = nsteps // 2
    slices = (nsteps + 1) // dsz + 1
    cube = np.zeros((n // 8, n // 8, slices), complex, order="F")
    O.cube_store(u0, cube, dsz, 0)
    ref = O.fft3d_run(g, csg, u0, 1, nsteps, absorber=ab, cube=cube, dsz=dsz)
    assert rel_l2(load(out, (n, n)), ref) < 1e-11
    assert rel_l2(load(out + ".cube", cube.shape), cube) < 1e-11
    if mode == "fft3d":
        _check_post3d(out, load(out, (n, n)), load(out + ".cube", cube.shape))
        # post::FarField::setPadding(CONSTANT), pad length 256, +-1 degree
        exp = O.far_field(load(out, (n, n)), None, 256, g.k, g.dx, g.dy, -1.0, 1.0, padding="constant")
        got = np.fromfile(out + ".ffconst", dtype=np.float64).reshape(exp.shape, order="F")
        assert rel_l2(got, exp) < 1e-11
        _check_saved(out, n, cube.shape, g.k)


def _check_saved(out, n, cshape, k):
    """ParaxialSimulation::save of the same run: each module's dataset in the
    reference's layout (matrices as {n_cols, n_rows}, cubes {n_rows, n_cols, n_slices}
    over the column-major buffer) holds the values the modules returned above, and the
    far field carries phiMin, phiMax, qmin, qmax (src/farFieldPost.cpp:131-139)."""
    import test_h5_save as H
    if not (os.path.exists(H.H5LIB) and os.path.exists(H.H5DUMP)):
        pytest.skip("HDF5 writer or h5dump not available")
    f = out + ".h5"
    tmp = pathlib.Path(out).parent
    _, blocks = H.header(f)
    assert H.dims(blocks["exitIntensity"]) == (n, n)
    np.testing.assert_array_equal(H.data(f, "/data/exitIntensity", np.float64, tmp).reshape((n, n)).T,
                                  _dbl(out + ".exitint", (n, n)))
    assert H.dims(blocks["exitPhase"]) == (30, 20)
    np.testing.assert_array_equal(H.data(f, "/data/exitPhase", np.float64, tmp).reshape((30, 20)).T,
                                  _dbl(out + ".exitphase_rs", (20, 30)))
    assert H.dims(blocks["amplitude"]) == cshape
    np.testing.assert_array_equal(H.data(f, "/data/amplitude", np.float64, tmp),
                                  np.fromfile(out + ".intcube", dtype=np.float64))
    assert "H5T_STD_U8LE" in blocks["LogAmplitude"] and H.dims(blocks["LogAmplitude"]) == cshape
    np.testing.assert_array_equal(H.data(f, "/data/LogAmplitude", np.uint8, tmp),
                                  np.fromfile(out + ".logu8", dtype=np.uint8))
    ff = np.fromfile(out + ".ffconst", dtype=np.float64)
    fshape = H.dims(blocks["farField"])
    np.testing.assert_array_equal(H.data(f, "/data/farField", np.float64, tmp), ff)
    for a, v in (("phiMin", -1.0), ("phiMax", 1.0), ("qmin", -k * np.pi / 180), ("qmax", k * np.pi / 180)):
        assert float(H.attr_value(f, "/data/farField/" + a)) == pytest.approx(v, rel=1e-5)
    assert H.attr_value(f, "/data/solver") == '"FFTSolver3D"'
    assert H.attr_value(f, "/data/description") == '"drop-in sphere"'
    assert len(fshape) == 2


def _dbl(path, shape):
    return np.fromfile(path, dtype=np.float64).reshape(shape, order="F")


def _check_post3d(out, last, cube):
    # post::ExitIntensity / ExitPhase (resized 20 x 30) / ExitField / Intensity / Phase /
    # IntensityUint8 / LogIntensityUint8 as the driver's solver returned them, against the
    # oracle's restatement on the same field and cube (last-ulp differences only)
    n = last.shape[0]
    for ext, q, shape, er, ec in ((".exitint", 1, (n, n), 0, 0), (".exitphase_rs", 2, (20, 30), 20, 30),
                                  (".exitfield", 3, (n, n), 0, 0)):
        exp = O.field_quantity(last, q, er, ec)
        assert np.max(np.abs(_dbl(out + ext, shape) - exp)) <= 4e-15 * np.max(np.abs(exp)), ext
    for ext, q in ((".intcube", 0), (".phasecube", 2)):
        exp = O.field_quantity(cube, q)
        assert np.max(np.abs(_dbl(out + ext, cube.shape) - exp)) <= 2e-15 * np.max(np.abs(exp)), ext
    for ext, exp in ((".u8", O.intensity_uint8(cube)),
                     (".logu8", O.log_intensity_uint8(cube, min_value=1e-3, max_value=0.9))):
        got = np.fromfile(out + ext, dtype=np.uint8).reshape(cube.shape, order="F")
        d = np.abs(got.astype(int) - exp.astype(int))
        assert d.max() <= 1 and np.count_nonzero(d) <= 1, ext


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["adi", "adi_reset"])
def test_dropin_adi(tmp_path, mode):
    n, nsteps = 64, 4
    out = run(tmp_path, mode, n, nsteps)
    g = _grid3d(n, nsteps)
    u0 = gaussian_field(g, n, 3000.0)
    csg = O.CSGMaterial(*spheres_scene([(0.0, 0.0, 0.0, 4000.0, 5e-6, 1e-7)]))
    ref = O.adi_run(g, csg, u0, 1, nsteps)
    assert rel_l2(load(out, (n, n)), ref) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("mode,nx", [("cn", 1024), ("cn_step", 300), ("cn_curved", 777), ("fft2d", 512),
                                     ("cn_eq", 300), ("fft2d", 500)])
def test_dropin_2d(tmp_path, mode, nx):
    nsteps, dz = 40, 500.0
    out = run(tmp_path, mode, nx, nsteps)
    if mode == "cn_curved":
        R, w = 5e6, 120.0
        L = nsteps * dz
        lo, hi = -L * L / (2 * R) - 2 * w - 200, w + 200
    else:
        lo, hi = -500.0, 500.0
    dx = (hi - lo) / (nx - 1)
    g = O.grid(lo, hi, dx, 0.0, nsteps * dz, dz, K_01)
    if mode == "cn_curved":
        mat = O.CurvedMaterial(R=R, width=w, clad_delta=4.5e-5, clad_beta=3.5e-6, dx=dx, dz=dz)
    else:
        mat = O.SlabMaterial(-50.0, 50.0, 0.0, 0.0, 4.5e-5, 3.5e-6)
    u0 = np.ones(nx, complex)
    sol = np.zeros((nx, nsteps + 1), complex, order="F")
    sol[:, 0] = u0
    if mode == "fft2d":
        ref = O.fft2d_run(g, mat, u0, 1, nsteps, solution=sol)
    else:
        # cn_eq: CrankNicholson::setEquation with a ParaxialEquation subclass whose
        # coefficients vary in x and z (the driver's PolyEquation)
        eq = O.PolyEq([[1.0, 1e-4, 1e-7, 0.0], [1.0, 2e-4, 0.0, 0.0], [1.0, -1e-4, 2e-8, 0.0],
                       [3e-6, 1e-9, 0.0, 0.0]]) if mode == "cn_eq" else None
        ref = O.cn_run(g, mat, u0, 1, nsteps, bc=O.BC_TRANSPARENT, solution=sol, equation=eq)
        if eq is not None:
            assert rel_l2(O.cn_run(g, mat, u0, 1, nsteps, bc=O.BC_TRANSPARENT), ref) > 1e-4
    assert rel_l2(load(out, (nx,)), ref) < 1e-10
    assert rel_l2(load(out + ".sol", sol.shape), sol) < 1e-10
    if mode == "cn":
        # post::Intensity, Phase (resized 40 x 7) of the solution and the exit quantities
        # as vectors, against the oracle's restatement on the driver's own field
        last, dsol = load(out, (nx,)), load(out + ".sol", sol.shape)
        for ext, src, q, shape, er, ec in ((".int2d", dsol, 0, sol.shape, 0, 0),
                                           (".phase2d_rs", dsol, 2, (40, 7), 40, 7),
                                           (".exitint", last, 1, (nx,), 0, 0),
                                           (".exitphase", last, 2, (nx,), 0, 0),
                                           (".exitfield", last, 3, (nx,), 0, 0)):
            exp = O.field_quantity(src, q, er, ec)
            got = np.fromfile(out + ext, dtype=np.float64).reshape(shape, order="F")
            assert np.max(np.abs(got - exp)) <= 2e-15 * np.max(np.abs(exp)), ext
    if mode == "cn_curved":
        # post::Transmittivity after every step, from the device-stored columns
        tr = np.fromfile(out + ".trans", dtype=np.float64)
        T, i0 = O.transmittivity_curved(sol, lo, dx, 0.0, dz, R, w)
        assert tr.size == nsteps + 1
        assert tr[0] == pytest.approx(i0, rel=1e-10)
        assert np.allclose(tr[1:], T, rtol=1e-10, atol=0)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(DRIVER), reason="drop-in driver not built")
@pytest.mark.parametrize("mode,prec,tol", [("scatter", 0, 1e-9), ("scatter_user", 0, 1e-9),
                                           ("scatter", 1, 3e-5), ("scatter_adi", 0, 1e-9)])
def test_dropin_generic_scattering_far_field(tmp_path, mode, prec, tol):
    # GenericScattering as Examples/sphere.cpp drives it (128 points per side, 32
    # steps): reference run, sphere run and far field, against the oracle flow
    n, nsteps, r = 128, 32, 500.0
    out = run(tmp_path, mode, n, nsteps, prec)
    xmin, dx = -1.5 * r, 3.0 * r / n
    zmin, zmax = -1.05 * r, 1.05 * r
    g = O.grid(xmin, xmin + (n - 1) * dx, dx, zmin, zmax, (zmax - zmin) / nsteps, 2 * np.pi / 0.1569)
    dz_ref = (zmax - zmin) / 3.0
    nz_ref = int(np.floor((zmax - zmin) / dz_ref + 1.0))
    g_ref = O.grid(g.xmin, g.xmax, g.dx, zmin, zmax, dz_ref, g.k)
    csg = spheres_scene([(0.0, 0.0, 0.0, r, 8.9e-6, 7e-7)])
    u0 = O.gauss3d_field(g, O.Gauss(g.k, 1.0, 0.0, 400.0 * r, 0.0, 0.0, 0.0, 0.0), n)
    prop = O.adi_run if mode == "scatter_adi" else O.fft3d_run  # the propagator the example selects
    ref = prop(g_ref, O.CSGMaterial(*csg, reference_run=True), u0, 1, nz_ref - 1)
    sol = prop(g, O.CSGMaterial(*csg), u0, 1, nsteps)
    exp = O.far_field(sol, ref, 1024, g.k, g.dx, g.dy, -0.05, 0.05, 512, 512)
    ff = np.fromfile(out, dtype=np.float64).reshape(exp.shape, order="F")
    err = rel_l2(ff, exp)
    announce(mode, "far field", ff.shape, "rel_l2", err)
    assert err < tol
