"""GPU parity of the split-step solvers on line lengths that are not a power of
two (Bluestein passes, paxpro_amd/csrc/pxg_bluestein.hip) against the CPU FP64
oracle, whose transforms of such lengths are the O(N^2) definition (pinned by
brute force in tests/test_oracle.py).  FFTW plans any length
(src/fftSolver3D.cpp:61-62, src/fftSolver2D.cpp:68-69) and the reference's
truncating node count makes 511-node grids ordinary
(src/paraxialSimulation.cpp:107-120).

Tolerances: FP64 1e-11 relative L2 (the Bluestein convolution adds a few
rounding steps: still ~1e-13); mixed precision (complex64 storage) 1e-5.
"""
import numpy as np
import pytest

import paxpro_amd as P
import pyoracle as O
from gpu_common import announce, rel_l2
from scenes import K_01, gaussian_field, random_spheres, spheres_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = P.Ctx(0)
    yield c
    c.close()


def _grid(n, dx, dy, zmin, dz, nsteps, k=K_01):
    xmin, ymin = -dx * (n // 2), -dy * (n // 2)
    return O.grid(xmin, xmin + (n - 1) * dx, dx, zmin, zmin + nsteps * dz, dz, k,
                  ymin=ymin, ymax=ymin + (n - 1) * dy, dy=dy)


def _prob(kind, g, n, nsteps, **kw):
    return P.problem(kind, n, nsteps + 1, g.dx, g.dz, g.k, xmin=g.xmin, ymin=g.ymin, dy=g.dy, zmin=g.zmin,
                     **kw)


def _gauss(g, n, w):
    x = g.xmin + g.dx * np.arange(n)
    y = g.ymin + g.dy * np.arange(n)
    return np.asfortranarray(np.outer(np.exp(-(y / w) ** 2), np.exp(-(x / w) ** 2)).astype(complex))


@pytest.mark.parametrize("n,nsteps,dy_ratio", [(100, 8, 1.0), (243, 5, 0.8), (511, 3, 1.0), (1000, 1, 1.1),
                                               (3, 4, 1.0), (7, 4, 1.3), (12, 6, 1.0)])
def test_fft3d_sphere_vs_oracle(ctx, n, nsteps, dy_ratio):
    announce("bluestein fft3d", n, nsteps, dy_ratio)
    ext = 20000.0
    dx = ext / n
    g = _grid(n, dx, dx * dy_ratio, -nsteps * 100.0, 200.0, nsteps)
    csg = spheres_scene(random_spheres(0x5EED0002, 8, (-6000, 6000), (-300, 300), (800, 3000), (1e-6, 1e-5)))
    ab = O.Absorber3D(min(5, n // 3), min(3, n // 3), 0.3, 0.2)
    u0 = _gauss(g, n, 4000.0)
    ref = O.fft3d_run(g, O.CSGMaterial(*csg), u0, 1, nsteps, absorber=ab)
    plan = P.Plan(ctx, _prob(P.FFT3D, g, n, nsteps, absorb=(ab.npix_x, ab.npix_y, ab.inv_x, ab.inv_y)))
    plan.material_csg(*csg)
    plan.set_field(u0)
    plan.step(1, nsteps)
    err = rel_l2(plan.get_field(), ref)
    announce("rel_l2", err)
    assert err < 1e-11
    plan.close()


def test_fft3d_mixed_and_batched_wavenumbers(ctx):
    """n = 300, three photon energies in one plan, complex64 storage: each field
    against the oracle run at its own wavenumber; the FP64 plan to 1e-11."""
    n, nsteps = 300, 6
    announce("bluestein fft3d batch", n)
    ks = np.array([K_01, K_01 * 1.3, K_01 * 0.7])
    dx = 15000.0 / n
    sph = random_spheres(0x5EED0005, 6, (-4000, 4000), (-200, 200), (600, 2500), (1e-6, 1e-5))
    csg = spheres_scene(sph)
    for prec, tol in ((P.PREC_FP64, 1e-11), (P.PREC_MIXED, 1e-5)):
        g0 = _grid(n, dx, dx, -300.0, 100.0, nsteps)
        u0 = _gauss(g0, n, 3000.0)
        plan = P.Plan(ctx, _prob(P.FFT3D, g0, n, nsteps, batch=3, precision=prec))
        plan.set_wavenumbers(ks)
        plan.material_csg(*csg)
        for b in range(3):
            plan.set_field(u0, b)
        plan.step(1, nsteps)
        for b in range(3):
            g = _grid(n, dx, dx, -300.0, 100.0, nsteps, k=ks[b])
            ref = O.fft3d_run(g, O.CSGMaterial(*csg), u0, 1, nsteps)
            err = rel_l2(plan.get_field(b), ref)
            announce("prec", prec, "field", b, "rel_l2", err)
            assert err < tol
        plan.close()


@pytest.mark.parametrize("p,q", [(3, 0), (100, 37), (255, 256)])
def test_fft3d_plane_wave_eigenmode_511(ctx, p, q):
    """Vacuum, n = 511, dy != dx: the grid plane wave comes back times
    exp(-i s dz (kx^2 + ky^2)/(2k)), kx = 2 pi min(p, n-p)/(dx n)."""
    n, s = 511, 4
    g = _grid(n, 40.0, 31.0, 0.0, 2000.0, s)
    r = np.arange(n)[:, None]
    c = np.arange(n)[None, :]
    u0 = np.asfortranarray(np.exp(2j * np.pi * (p * c + q * r) / n))
    plan = P.Plan(ctx, _prob(P.FFT3D, g, n, s))
    plan.set_field(u0)
    plan.step(1, s)
    kx = 2 * np.pi * min(p, n - p) / (g.dx * n)
    ky = 2 * np.pi * min(q, n - q) / (g.dy * n)
    lam = np.exp(-1j * s * g.dz * (kx * kx + ky * ky) / (2 * g.k))
    err = rel_l2(plan.get_field(), lam * u0)
    announce("eigenmode 511", p, q, err)
    assert err < 1e-12
    plan.close()


def test_fft3d_split_calls_cube_and_graphs(ctx):
    """Step-by-step calls equal one call; the block-averaged cube (120/8 = 15
    cells per side) and graph replays equal the stream enqueue."""
    n, nsteps = 120, 6
    g = _grid(n, 100.0, 100.0, -300.0, 100.0, nsteps)
    csg = spheres_scene([(0.0, 0.0, 0.0, 3000.0, 5e-6, 1e-7)])
    u0 = _gauss(g, n, 3000.0)
    cube = np.zeros((15, 15, nsteps // 2 + 1), complex, order="F")
    O.cube_store(u0, cube, 2, 0)
    ref = O.fft3d_run(g, O.CSGMaterial(*csg), u0, 1, nsteps, cube=cube, dsz=2)
    plan = P.Plan(ctx, _prob(P.FFT3D, g, n, nsteps, store_dsx=8, store_dsz=2))
    plan.material_csg(*csg)
    plan.set_field(u0)
    for st in range(1, nsteps + 1):
        plan.step(st, 1)
    assert rel_l2(plan.get_field(), ref) < 1e-11
    assert rel_l2(plan.get_solution(), cube) < 1e-11
    plan.close()
    outs = []
    for graphs in (False, True):
        plan = P.Plan(ctx, _prob(P.FFT3D, g, n, nsteps))
        plan.material_csg(*csg)
        plan.set_graphs(graphs)
        for rep in range(2):
            plan.set_field(u0)
            plan.step(1, nsteps)
            outs.append(plan.get_field())
        plan.close()
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


@pytest.mark.parametrize("n,material", [(5, "csg"), (100, "curved"), (511, "csg"), (1000, "curved"),
                                        (4000, "csg")])
def test_fft2d_vs_oracle(ctx, n, material):
    nsteps, dz = 12, 400.0
    announce("bluestein fft2d", n, material)
    if material == "curved":
        R, w = 5e6, 120.0
        L = nsteps * dz
        lo, hi = -L * L / (2 * R) - 2 * w - 200, w + 200
        dx = (hi - lo) / (n - 1)
        mat = O.CurvedMaterial(R=R, width=w, clad_delta=4.5e-5, clad_beta=3.5e-6, dx=dx, dz=dz)
    else:
        lo, dx = -1000.0, 2000.0 / (n - 1)
        csg = spheres_scene([(0.0, 0.0, 2000.0, 600.0, 8e-6, 2e-7), (300.0, 0.0, 3000.0, 300.0, 3e-6, 1e-7)])
        mat = O.CSGMaterial(*csg)
    g = O.grid(lo, lo + (n - 1) * dx, dx, 0.0, nsteps * dz, dz, K_01)
    x = lo + dx * np.arange(n)
    u0 = (1.0 + 0.3 * np.exp(-((x - x.mean()) / (0.2 * (x[-1] - x[0]) + 1.0)) ** 2)).astype(complex)
    sol = np.zeros((n, nsteps + 1), complex, order="F")
    sol[:, 0] = u0
    ref = O.fft2d_run(g, mat, u0, 1, nsteps, solution=sol)
    plan = P.Plan(ctx, P.problem(P.FFT2D, n, nsteps + 1, dx, dz, K_01, xmin=lo, zmin=0.0, store_dsx=1,
                                 store_dsz=1))
    if material == "curved":
        plan.material_curved(R=R, width=w, clad_delta=4.5e-5, clad_beta=3.5e-6, dx=dx, dz=dz)
    else:
        plan.material_csg(*csg)
    plan.set_field(u0)
    plan.step(1, nsteps)
    err = rel_l2(plan.get_field(), ref)
    announce("rel_l2", err)
    assert err < 1e-11
    assert rel_l2(plan.get_solution(), sol) < 1e-11
    plan.close()


def test_work_group_shapes_and_limits(ctx):
    g = _grid(300, 50.0, 50.0, 0.0, 10.0, 2)
    plan = P.Plan(ctx, _prob(P.FFT3D, g, 300, 2))
    assert plan.lines() == (1, 1)
    plan.autotune()                     # one shape: a no-op
    assert plan.lines() == (1, 1)
    with pytest.raises(P.PxgError) as e:
        plan.set_lines(2, 1)
    assert e.value.code == P.ERR_UNSUPPORTED
    plan.close()
    with pytest.raises(P.PxgError) as e:
        P.plan_bytes(P.problem(P.FFT3D, 4097, 3, 1.0, 1.0, K_01))
    assert e.value.code == P.ERR_UNSUPPORTED
