"""GPU parity on the reference's own CSG scenes and on grids with dy != dx.

Scenes: Examples/primitiveGeomExample.cpp:15-33 (rotated box, cylinder carved
out by DIFFERENCE, a part rotated and moved after its shapes were added, an
ellipsoid from a scaled sphere), Examples/slab.cpp:14-27 (a tilted box slab),
and a two-module scene with a DIFFERENCE part and a cone (tests/scenes.py).
Each runs through FFT3D, ADI and FFT2D against the CPU FP64 oracle, on grids
whose y spacing differs from x, so the kx-from-columns / ky-from-rows
orientation of the propagator (src/fftSolver3D.cpp:24-50,103-115) and the
dx/dy terms of both ADI passes (src/alternatingDirectionSolver.cpp:38-51,
103-116) are checked on the device.

Tolerances as elsewhere: FP64 1e-11 (split step) and 1e-10 (ADI) relative L2;
mixed precision 1e-5.
"""
import numpy as np
import pytest

import paxpro_amd as P
import pyoracle as O
from gpu_common import announce, rel_l2
from scenes import K_01, multi_module_scene, primitive_geom_scene, slab_example_scene

pytestmark = pytest.mark.gpu

SCENES = {"primitive": primitive_geom_scene, "multi": multi_module_scene}


@pytest.fixture(scope="module")
def ctx():
    c = P.Ctx(0)
    yield c
    c.close()


def _grid(n, dx, dy, zmin, dz, nsteps, k=K_01):
    xmin, ymin = -dx * n / 2, -dy * n / 2
    return O.grid(xmin, xmin + (n - 1) * dx, dx, zmin, zmin + nsteps * dz, dz, k,
                  ymin=ymin, ymax=ymin + (n - 1) * dy, dy=dy)


def _gauss(g, n, wx, wy, cx=0.0, cy=0.0):
    x = g.xmin + g.dx * np.arange(n)
    y = g.ymin + g.dy * np.arange(n)
    return np.asfortranarray(np.outer(np.exp(-((y - cy) / wy) ** 2),
                                      np.exp(-((x - cx) / wx) ** 2)).astype(complex))


def _plan(ctx, kind, g, n, nsteps, prec=P.PREC_FP64, **kw):
    return P.Plan(ctx, P.problem(kind, n, nsteps + 1, g.dx, g.dz, g.k, xmin=g.xmin, ymin=g.ymin,
                                 dy=g.dy, zmin=g.zmin, precision=prec, **kw))


def _scene_fitting(scene, ex, ey):
    """The scene with its length unit s chosen to fit an ex x ey window: the
    primitive scene spans about -60s..35s in x and +-15s in y, the two-module
    scene about +-4.5s in both."""
    if scene == "primitive":
        return primitive_geom_scene(s=min(ex / 130.0, ey / 32.0))
    return multi_module_scene(s=min(ex, ey) / 10.0)


# ------------------------------------------------------------------ FFT3D
@pytest.mark.parametrize("scene", sorted(SCENES))
@pytest.mark.parametrize("prec,tol", [(P.PREC_FP64, 1e-11), (P.PREC_MIXED, 1e-5)])
def test_fft3d_reference_scenes_anisotropic(ctx, scene, prec, tol):
    n, nsteps = 128, 40
    announce("fft3d scene", scene, prec, n, nsteps)
    g = _grid(n, 125.0, 70.0, -2400.0, 120.0, nsteps)
    csg = SCENES[scene]()
    u0 = _gauss(g, n, 4500.0, 2500.0, cx=-500.0)
    ref = O.fft3d_run(g, O.CSGMaterial(*csg), u0, 1, nsteps)
    plan = _plan(ctx, P.FFT3D, g, n, nsteps, prec)
    plan.material_csg(*csg)
    plan.set_field(u0)
    plan.step(1, nsteps)
    err = rel_l2(plan.get_field(), ref)
    announce("rel_l2", err)
    assert err < tol
    # the scene really acted on the field
    vac = O.fft3d_run(g, O.NoMaterial, u0, 1, nsteps)
    assert rel_l2(vac, ref) > 1e-4
    plan.close()


def test_fft3d_slab_example(ctx):
    """Examples/slab.cpp: 250 nm window, lambda 0.1569 nm, steps that cross the
    slab's end face at z = 2500 nm (border cells through the crossing integral)."""
    n, nsteps, first = 128, 16, 120
    announce("fft3d slab example", n, first, nsteps)
    dx = 250.0 / n
    k = 2 * np.pi / 0.1569
    g = O.grid(-125.0, -125.0 + (n - 1) * dx, dx, 0.0, 5000.0, 5000.0 / 256, k)
    csg = slab_example_scene()
    u0 = _gauss(g, n, 20.0, 20.0, cy=-20.0)
    ref = O.fft3d_run(g, O.CSGMaterial(*csg), u0, first, nsteps)
    plan = P.Plan(ctx, P.problem(P.FFT3D, n, 257, dx, g.dz, k, xmin=-125.0, zmin=0.0))
    plan.material_csg(*csg)
    plan.set_field(u0)
    plan.step(first, nsteps)
    err = rel_l2(plan.get_field(), ref)
    announce("rel_l2", err)
    assert err < 1e-11
    plan.close()


@pytest.mark.parametrize("p,q", [(3, 0), (0, 5), (7, 2), (40, 100)])
@pytest.mark.parametrize("prec,tol", [(P.PREC_FP64, 1e-12), (P.PREC_MIXED, 1e-6)])
def test_fft3d_plane_wave_eigenmode_on_device(ctx, p, q, prec, tol):
    """Vacuum, dy != dx: the grid plane wave exp(2 pi i (p c + q r)/N) comes back
    times exp(-i s dz (kx^2 + ky^2)/(2k)), kx from (p, dx), ky from (q, dy)."""
    n, s = 128, 5
    announce("fft3d eigenmode", p, q, prec)
    g = _grid(n, 39.0625, 23.0, 0.0, 3000.0, s)
    r = np.arange(n)[:, None]
    c = np.arange(n)[None, :]
    u0 = np.asfortranarray(np.exp(2j * np.pi * (p * c + q * r) / n))
    plan = _plan(ctx, P.FFT3D, g, n, s, prec)
    plan.set_field(u0)
    plan.step(1, s)
    kx = 2 * np.pi * min(p, n - p) / (g.dx * n)
    ky = 2 * np.pi * min(q, n - q) / (g.dy * n)
    lam = np.exp(-1j * s * g.dz * (kx * kx + ky * ky) / (2 * g.k))
    err = rel_l2(plan.get_field(), lam * u0)
    announce("rel_l2", err, "phase", np.angle(lam))
    assert err < tol
    plan.close()


# ------------------------------------------------------------------ ADI
@pytest.mark.parametrize("scene", sorted(SCENES))
@pytest.mark.parametrize("quirks", [0, P.FIX_ADI_Y_DIAG])
@pytest.mark.parametrize("dx,dy", [(80.0, 55.0), (80.0, 2.5)])
def test_adi_reference_scenes_anisotropic(ctx, scene, quirks, dx, dy):
    """dx = 80, dy = 55 nm: both passes on the tile-Jacobi path; dy = 2.5 nm:
    the y-implicit pass is not dominant enough and runs the sequential sweep
    while the x-implicit pass stays on tiles.  200 cells: partial tiles."""
    n, nsteps = 200, 3
    announce("adi scene", scene, quirks, dx, dy)
    g = _grid(n, dx, dy, -300.0, 150.0, nsteps)   # planes z = -150 .. 225 cut the scene
    csg = _scene_fitting(scene, n * dx, n * dy)
    wy = 0.3 * n * dy
    u0 = _gauss(g, n, 0.3 * n * dx, wy)
    for use_tbc in (True, False):
        ref = O.adi_run(g, O.CSGMaterial(*csg), u0, 1, nsteps, use_tbc=use_tbc, quirks=quirks)
        plan = _plan(ctx, P.ADI, g, n, nsteps, use_tbc=int(use_tbc), quirks=quirks)
        plan.material_csg(*csg)
        plan.set_field(u0)
        plan.step(1, nsteps)
        err = rel_l2(plan.get_field(), ref)
        announce("adi rel_l2", use_tbc, err)
        assert err < 1e-10
        plan.close()


@pytest.mark.parametrize("mx,my", [(1, 1), (5, 9), (30, 2)])
@pytest.mark.parametrize("dy", [55.0, 2.5])
def test_adi_sine_modes_on_device(ctx, mx, my, dy):
    """No transparent boundary, vacuum, dy != dx: a product of sine modes is
    multiplied per step by the product of the two half-step factors."""
    n, dx, dz = 63, 80.0, 150.0
    announce("adi sine modes", mx, my, dy)
    g = _grid(n, dx, dy, 0.0, dz, 1)
    r = np.arange(n)[:, None]
    c = np.arange(n)[None, :]
    tx, ty = np.pi * mx / (n + 1), np.pi * my / (n + 1)
    u0 = np.asfortranarray((np.sin(ty * (r + 1)) * np.sin(tx * (c + 1))).astype(complex))
    plan = _plan(ctx, P.ADI, g, n, 1, use_tbc=0)
    plan.set_field(u0)
    plan.step(1, 1)
    h, k = dz / 2, g.k
    ax, ay = 1j / (k * dx * dx), 1j / (k * dy * dy)
    fx = (1 / h - ay * (1 - np.cos(ty))) / (1 / h + ax * (1 - np.cos(tx)))
    fy = (1 / h - ax * (1 - np.cos(tx))) / (1 / h + ay * (1 - np.cos(ty)))
    err = rel_l2(plan.get_field(), fx * fy * u0)
    announce("rel_l2", err)
    assert err < 1e-12
    plan.close()


# ------------------------------------------------------------------ FFT2D
@pytest.mark.parametrize("scene", ["primitive", "multi"])
def test_fft2d_reference_scenes(ctx, scene):
    """The 2D solvers sample the scene at y = 0 (src/paraxialSimulation.cpp:513-536).
    (The slab example lies at y in [-61.5, -11.5] nm, so its y = 0 slice is empty.)"""
    n, nsteps = 512, 30
    announce("fft2d scene", scene)
    csg = SCENES[scene]()
    dx, zmin, dz, w, cx = 16000.0 / n, -2400.0, 160.0, 3000.0, -500.0
    xmin = -dx * n / 2
    g = O.grid(xmin, xmin + (n - 1) * dx, dx, zmin, zmin + nsteps * dz, dz, K_01)
    x = xmin + dx * np.arange(n)
    u0 = np.exp(-((x - cx) / w) ** 2).astype(complex)
    ref = O.fft2d_run(g, O.CSGMaterial(*csg), u0, 1, nsteps)
    plan = P.Plan(ctx, P.problem(P.FFT2D, n, nsteps + 1, dx, dz, K_01, xmin=xmin, zmin=zmin))
    plan.material_csg(*csg)
    plan.set_field(u0)
    plan.step(1, nsteps)
    err = rel_l2(plan.get_field(), ref)
    announce("rel_l2", err)
    assert err < 1e-11
    vac = O.fft2d_run(g, O.NoMaterial, u0, 1, nsteps)
    assert rel_l2(vac, ref) > 1e-5
    plan.close()
