"""GPU parity of the far field (post::FarField::result, src/farFieldPost.cpp:141-230)
and of the GenericScattering flow around it (src/genericScattering.cpp:148-213):
a vacuum reference run, the run with the material, the far field of their
difference, against the CPU oracle through the C ABI.

Tolerances: the far field of a given field is FP64 from the field on, so the GPU
and the oracle agree to 1e-12 relative L2 (different, equally accurate FFT
roundings); the full scattering flow in FP64 to 1e-9 (the two propagations carry
their own 1e-12-level differences into the subtraction).
"""
import numpy as np
import pytest

import paxpro_amd as P
import pyoracle as O
from gpu_common import announce, fft3d_problem, rel_l2
from scenes import spheres_scene

pytestmark = pytest.mark.gpu

LAM = 0.1569
K = 2 * np.pi / LAM


@pytest.fixture(scope="module")
def ctx():
    c = P.Ctx(0)
    yield c
    c.close()


def _grid(n, r=500.0, nsteps=32):
    # the reference's sphere example (Examples/sphere.cpp) at n points per side
    xmin, ext = -1.5 * r, 3.0 * r
    dx = ext / n
    zmin, zmax = -1.05 * r, 1.05 * r
    return O.grid(xmin, xmin + (n - 1) * dx, dx, zmin, zmax, (zmax - zmin) / nsteps, K)


def _random_field(n, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))


@pytest.mark.parametrize("kind,n,prec,pad,phi,export", [
    (P.FFT3D, 64, P.PREC_FP64, 64, 90.0, (0, 0)),         # M = n: nothing is padded
    (P.FFT3D, 64, P.PREC_FP64, 256, 90.0, (0, 0)),        # whole spectrum, M = 4 NP
    (P.FFT3D, 64, P.PREC_FP64, 1024, 1.0, (0, 0)),
    (P.FFT3D, 128, P.PREC_MIXED, 2048, 0.5, (16, 12)),
    (P.ADI, 48, P.PREC_FP64, 256, 2.0, (0, 0)),           # n below NP = 64
    (P.FFT3D, 64, P.PREC_FP64, 8192, 0.3, (0, 0)),        # linear fit: 4064-entry segments
    (P.FFT3D, 100, P.PREC_FP64, 333, 3.0, (0, 0)),        # odd signal length, direct sums
    (P.ADI, 50, P.PREC_FP64, 3000, 1.0, (0, 0)),          # even M, 1475-entry segments
])
@pytest.mark.parametrize("padding", ["constant", "linear", "linear_signed"])
def test_far_field_padding_vs_oracle(ctx, kind, n, prec, pad, phi, export, padding):
    # setPadding(CONSTANT / LINEAR_FIT): the device adds the padded segments' spectra in
    # closed form (geometric and arithmetico-geometric sums); the oracle pads every line
    # explicitly, as the reference does, and transforms M points
    announce("far field padding", padding, kind, n, prec, pad, phi, export)
    # the linear fit's first segment wraps its abscissa near 2^64 as the reference does
    # (runs of a 2048-step staircase beyond 1024 entries); "linear_signed" sets the
    # plan's PXG_FIX_LINEAR_PAD
    mode = P.PAD_CONSTANT if padding == "constant" else P.PAD_LINEAR_FIT
    quirks = P.FIX_LINEAR_PAD if padding == "linear_signed" else 0
    g = _grid(n)
    prob = P.problem(kind, n, 3, g.dx, g.dz, K, xmin=g.xmin, ymin=g.ymin, dy=g.dy, zmin=g.zmin,
                     precision=prec, quirks=quirks)
    plan, refp = P.Plan(ctx, prob), P.Plan(ctx, prob)
    plan.set_field(_random_field(n, 3) + 2.0)  # a mean that the padding carries
    refp.set_field(_random_field(n, 4))
    sol, ref = plan.get_field(), refp.get_field()
    out = plan.far_field(0, refp, 0, pad, -phi, phi, *export, padding=mode)
    exp = O.far_field(sol, ref, pad, K, g.dx, g.dy, -phi, phi, *export, padding=padding.replace("_signed", ""),
                      xmin=g.xmin, quirks=quirks)
    assert out.shape == exp.shape
    err = rel_l2(out, exp)
    announce("shape", out.shape, "rel_l2", err)
    assert err < 1e-11
    zero = plan.far_field(0, refp, 0, pad, -phi, phi, *export)
    if pad > n:
        assert rel_l2(zero, exp) > 1e-3  # the padding changes the pattern
    plan.close()
    refp.close()


@pytest.mark.parametrize("kind,n,prec,pad,phi,export", [
    (P.FFT3D, 64, P.PREC_FP64, 64, 90.0, (0, 0)),         # whole spectrum, M = n
    (P.FFT3D, 64, P.PREC_FP64, 1024, 1.0, (0, 0)),        # M = 16 NP: residue split
    (P.FFT3D, 128, P.PREC_MIXED, 2048, 0.5, (16, 12)),    # complex64 fields, export resize
    (P.ADI, 48, P.PREC_FP64, 256, 2.0, (0, 0)),           # n below NP = 64, zero-filled
    (P.FFT3D, 512, P.PREC_FP64, 32768, 0.05, (512, 512)), # the Python example's pad length
    (P.FFT3D, 100, P.PREC_FP64, 0, 90.0, (0, 0)),         # M = n = 100: direct window sums
    (P.FFT3D, 99, P.PREC_MIXED, 0, 90.0, (20, 30)),       # odd M: fftshift leaves the last entry
    (P.ADI, 64, P.PREC_FP64, 1000, 3.0, (0, 0)),          # pad length not a power of two
])
def test_far_field_of_given_fields_vs_oracle(ctx, kind, n, prec, pad, phi, export):
    announce("far field", kind, n, prec, pad, phi, export)
    g = _grid(n)
    prob = P.problem(kind, n, 3, g.dx, g.dz, K, xmin=g.xmin, ymin=g.ymin, dy=g.dy, zmin=g.zmin,
                     precision=prec)
    plan, refp = P.Plan(ctx, prob), P.Plan(ctx, prob)
    plan.set_field(_random_field(n, 1))
    refp.set_field(_random_field(n, 2))
    sol, ref = plan.get_field(), refp.get_field()  # as stored (complex64 for mixed)
    out = plan.far_field(0, refp, 0, pad, -phi, phi, *export)
    exp = O.far_field(sol, ref, pad, K, g.dx, g.dy, -phi, phi, *export)
    assert out.shape == exp.shape
    err = rel_l2(out, exp)
    announce("shape", out.shape, "rel_l2", err)
    assert err < 1e-12
    plan.close()
    refp.close()


@pytest.mark.parametrize("prec,tol", [(P.PREC_FP64, 1e-9), (P.PREC_MIXED, 3e-5)])
def test_generic_scattering_sphere_vs_oracle(ctx, prec, tol):
    # Examples/sphere.cpp at 128 points per side: reference run over 3 steps
    # without material, then the sphere run, far field within 0.05 degrees
    n, nsteps, r = 128, 32, 500.0
    announce("generic scattering sphere", n, nsteps, prec)
    g = _grid(n, r, nsteps)
    dz_ref = (g.zmax - g.zmin) / 3.0
    nz_ref = int(np.floor((g.zmax - g.zmin) / dz_ref + 1.0))  # truncating node count
    g_ref = O.grid(g.xmin, g.xmax, g.dx, g.zmin, g.zmax, dz_ref, K)
    csg = spheres_scene([(0.0, 0.0, 0.0, r, 8.9e-6, 7e-7)])
    gb = O.Gauss(K, 1.0, 0.0, 400.0 * r, 0.0, 0.0, 0.0, 0.0)
    u0 = O.gauss3d_field(g, gb, n)

    ref_o = O.fft3d_run(g_ref, O.CSGMaterial(*csg, reference_run=True), u0, 1, nz_ref - 1)
    sol_o = O.fft3d_run(g, O.CSGMaterial(*csg), u0, 1, nsteps)
    exp = O.far_field(sol_o, ref_o, 1024, K, g.dx, g.dy, -0.05, 0.05, 512, 512)

    vac = P.Plan(ctx, fft3d_problem(g_ref, n, nz_ref, precision=prec))
    vac.material_csg(*csg, reference_run=True)
    vac.set_field(u0)
    vac.step(1, nz_ref - 1)
    run = P.Plan(ctx, fft3d_problem(g, n, nsteps + 1, precision=prec))
    run.material_csg(*csg)
    run.set_field(u0)
    run.step(1, nsteps)
    out = run.far_field(0, vac, 0, 1024, -0.05, 0.05, 512, 512)
    err = rel_l2(out, exp)
    announce("far field", out.shape, "rel_l2", err, "peak", float(out.max()))
    assert out.shape == exp.shape
    assert err < tol
    vac.close()
    run.close()


@pytest.mark.parametrize("pad,phi,export", [
    (0, (-0.05, 0.12), (0, 0)),       # M = n, asymmetric interior window: rows != cols
    (512, (-0.3, 0.08), (20, 1000)),  # one export size above the matrix: no resize
    (4096, (-0.02, 0.03), (9, 7)),    # D = 32, small export
])
def test_far_field_anisotropic_grid(ctx, pad, phi, export):
    # dy != dx: the row window follows dy, the column window dx
    n = 128
    g = _grid(n)
    dy = 0.7 * g.dx
    prob = P.problem(P.FFT3D, n, 3, g.dx, g.dz, K, xmin=g.xmin, ymin=g.ymin, dy=dy, zmin=g.zmin)
    plan, refp = P.Plan(ctx, prob), P.Plan(ctx, prob)
    plan.set_field(_random_field(n, 7))
    refp.set_field(_random_field(n, 8))
    out = plan.far_field(0, refp, 0, pad, phi[0], phi[1], *export)
    exp = O.far_field(plan.get_field(), refp.get_field(), pad, K, g.dx, dy, phi[0], phi[1], *export)
    announce("anisotropic far field", pad, phi, export, out.shape)
    assert out.shape == exp.shape
    assert rel_l2(out, exp) < 1e-12
    plan.close()
    refp.close()


def test_far_field_errors(ctx):
    g = _grid(64)
    plan = P.Plan(ctx, fft3d_problem(g, 64, 3))
    with pytest.raises(P.PxgError) as e:
        plan.far_field(0, None, 0, 1024)  # field not set yet
    assert e.value.code == 5
    plan.set_field(_random_field(64, 3))
    with pytest.raises(P.PxgError) as e:
        plan.far_field(0, None, 0, (1 << 24) + 1)  # signal length beyond 2^24
    assert e.value.code == 2
    with pytest.raises(P.PxgError) as e:
        plan.far_field(0, None, 0, 1024, 1.0, -1.0)  # empty window
    assert e.value.code == 1
    p2 = P.Plan(ctx, P.problem(P.FFT2D, 64, 3, 1.0, 1.0, K))
    p2.set_field(np.ones(64, complex))
    with pytest.raises(P.PxgError) as e:
        p2.far_field(0, None, 0, 1024)
    assert e.value.code == 1
    # the plan as its own reference: the difference, and the far field, vanish
    assert not np.any(plan.far_field(0, plan, 0, 256, -3.0, 3.0))
    with pytest.raises(P.PxgError) as e:
        plan.far_field(0, plan, 1, 256)  # reference field index out of range
    assert e.value.code == 1
    # no reference: the plain far field of the field
    f = plan.get_field()
    out = plan.far_field(0, None, 0, 256, -3.0, 3.0)
    assert rel_l2(out, O.far_field(f, None, 256, K, g.dx, g.dy, -3.0, 3.0)) < 1e-12
    plan.close()
    p2.close()
