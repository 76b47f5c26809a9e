"""GPU parity of the per-step transmittivity sums (pxg_intensity_trapezoid, the
trapezoid of post::Transmittivity, src/transmittivity.cpp:30-83) on job plans of
curved guides, against the oracle's restatement applied to the same stored
solution columns (FP64 reduction order differs: 1e-13 relative)."""
import numpy as np
import pytest

import paxpro_amd as P
import pyoracle as O
from gpu_common import announce

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = P.Ctx(0)
    yield c
    c.close()


def _jobs():
    # (nx, lambda, R, width): uneven sizes as in the batched curved-guide workload
    return [(300, 0.1, 5e6, 120.0), (777, 0.07, 2e6, 80.0), (1500, 0.15, 2e7, 200.0)]


@pytest.mark.parametrize("prec", [P.PREC_FP64, P.PREC_MIXED])
def test_transmittivity_sums_on_job_plan(ctx, prec):
    nsteps, dz, L = 50, 100.0, 5000.0
    probs, geo = [], []
    for nx, lam, R, w in _jobs():
        lo, hi = -L * L / (2 * R) - 2 * w - 200, w + 200
        dx = (hi - lo) / (nx - 1)
        probs.append(P.problem(P.CN, nx, nsteps + 1, dx, dz, 2 * np.pi / lam, xmin=lo, precision=prec,
                               bc=P.BC_TRANSPARENT, store_dsx=1, store_dsz=1))
        geo.append((lo, dx, R, w))
    plan = P.JobPlan(ctx, probs)
    for b, (nx, lam, R, w) in enumerate(_jobs()):
        plan.material_curved_job(b, R=R, width=w, clad_delta=4.5e-5, clad_beta=3.5e-6, dx=probs[b].dx, dz=dz)
        plan.set_field(np.ones(nx, complex), b)
    plan.step(1, nsteps)
    for b, (lo, dx, R, w) in enumerate(geo):
        sol = plan.get_solution(b)
        steps, s, e = [1], [O.closest_index(0.0, lo, dx)], [O.closest_index(w, lo, dx)]
        for n in range(1, nsteps + 1):
            z = n * dz
            xs = -0.5 * z * z / R
            steps.append(n)
            s.append(O.closest_index(xs, lo, dx))
            e.append(O.closest_index(xs + w, lo, dx))
        sums = plan.intensity_trapezoid(steps, s, e, b)
        exp = np.array([O.trapezoid_intensity(sol[:, st], a, c, lo, dx) for st, a, c in zip(steps, s, e)])
        scale = np.array([((lo + c * dx) - (lo + a * dx)) / (c - a) * 0.5 for a, c in zip(s, e)])
        err = np.max(np.abs(sums * scale - exp) / np.abs(exp))
        announce("trapezoid job", b, prec, err)
        assert err < 1e-13
        T, _ = O.transmittivity_curved(sol, lo, dx, 0.0, dz, R, w)
        assert np.allclose(sums[1:] * scale[1:] / (sums[0] * scale[0]), T, rtol=1e-13, atol=0)


def test_trapezoid_errors(ctx):
    p = P.Plan(ctx, P.problem(P.CN, 64, 5, 1.0, 1.0, 60.0))  # no stored solution
    p.set_field(np.ones(64, complex))
    with pytest.raises(P.PxgError) as e:
        p.intensity_trapezoid([1], [0], [10])
    assert e.value.code == 5
    q = P.Plan(ctx, P.problem(P.CN, 64, 5, 1.0, 1.0, 60.0, store_dsx=1, store_dsz=1))
    q.set_field(np.ones(64, complex))
    q.step(1, 4)
    with pytest.raises(P.PxgError) as e:
        q.intensity_trapezoid([1], [0], [64])
    assert e.value.code == 1
    with pytest.raises(P.PxgError) as e:
        q.intensity_trapezoid([5], [0], [10])
    assert e.value.code == 1
    # a uniform field of ones: the sum is 2 (e - s)
    assert q.intensity_trapezoid([0], [3], [13])[0] == pytest.approx(20.0, rel=1e-15)
    p.close()
    q.close()


# ---- field quantities of the post-processing modules (pxg_field_quantity, pxg_field_uint8)
# The device result is checked against the oracle's restatement applied to the same
# device data (the field or stored solution as downloaded: differences are last-ulp
# ones of hypot / atan2 / log and of the summation order of resizeMatrix), and the
# quantity of the final field against the oracle's own run.

def _close(dev, exp, tol):
    assert dev.shape == exp.shape
    scale = max(float(np.max(np.abs(exp))), 1e-300)
    err = float(np.max(np.abs(dev - exp))) / scale
    assert err <= tol, err
    return err


def _u8_equal(dev, exp):
    # equal bytes, but for values whose unrounded image lies within an ulp of an integer
    assert dev.shape == exp.shape and dev.dtype == np.uint8
    d = np.abs(dev.astype(int) - exp.astype(int))
    assert d.max() <= 1 and np.count_nonzero(d) <= max(1, d.size // 100000), np.count_nonzero(d)


@pytest.mark.parametrize("prec", [P.PREC_FP64, P.PREC_MIXED])
def test_field_quantities_3d(ctx, prec):
    from gpu_common import fft3d_problem, rel_l2
    from scenes import fft3d_case, gaussian_field, spheres_scene
    n, nsteps, ds, dsz = 64, 12, 8, 4
    announce("field quantities 3d", n, nsteps, prec)
    g = fft3d_case(n, dz=10.0, nsteps=nsteps)
    u0 = gaussian_field(g, n, 3000.0)
    csg = spheres_scene([(0.0, 0.0, 0.0, 4000.0, 5e-6, 1e-7)])
    ref = O.fft3d_run(g, O.CSGMaterial(*csg), u0, 1, nsteps)
    plan = P.Plan(ctx, fft3d_problem(g, n, nsteps + 1, precision=prec, store=(ds, dsz)))
    plan.material_csg(*csg)
    plan.set_field(u0)
    plan.step(1, nsteps)
    u = plan.get_field()
    for q in (P.Q_ABS, P.Q_ABS2, P.Q_ARG, P.Q_REAL):
        dev = plan.field_quantity(q)
        assert dev.shape == (n, n, 1)
        _close(dev[:, :, 0], O.field_quantity(u, q), 1e-15)
        rs = plan.field_quantity(q, export_rows=20, export_cols=30)
        assert rs.shape == (20, 30, 1)
        _close(rs[:, :, 0], O.field_quantity(u, q, 20, 30), 4e-15)
    # an export size reaching the field's keeps it whole (resizeMatrix's early return)
    assert plan.field_quantity(P.Q_ABS, export_rows=64, export_cols=30).shape == (n, n, 1)
    tol = 1e-11 if prec == P.PREC_FP64 else 3e-5
    assert rel_l2(plan.field_quantity(P.Q_ABS2)[:, :, 0], O.field_quantity(ref, P.Q_ABS2)) < tol
    # the block-averaged cube (getSolution3D): Intensity, Phase, the 8-bit images
    cube = plan.get_solution()
    for q in (P.Q_ABS, P.Q_ARG):
        dev = plan.field_quantity(q, source=P.SRC_SOLUTION, export_rows=2, export_cols=2)
        assert dev.shape == cube.shape  # cubes are never resized
        _close(dev, O.field_quantity(cube, q), 1e-15)
    _u8_equal(plan.field_uint8(), O.intensity_uint8(cube))
    _u8_equal(plan.field_uint8(log_scale=True), O.log_intensity_uint8(cube))
    a = np.abs(cube)
    lo, hi = float(np.quantile(a, 0.2)), float(np.quantile(a, 0.9))
    for mn, mx in ((lo, None), (None, hi), (lo, hi), (a.min() / 2, a.max() * 2)):
        _u8_equal(plan.field_uint8(log_scale=True, min_value=mn, max_value=mx),
                  O.log_intensity_uint8(cube, min_value=mn, max_value=mx))
    # IntensityUint8 of the current field
    _u8_equal(plan.field_uint8(source=P.SRC_FIELD)[:, :, 0], O.intensity_uint8(u))
    plan.close()


@pytest.mark.parametrize("prec", [P.PREC_FP64, P.PREC_MIXED])
def test_field_quantities_2d(ctx, prec):
    nx, nsteps, dz = 300, 30, 500.0
    announce("field quantities 2d", nx, nsteps, prec)
    dx = 1000.0 / (nx - 1)
    plan = P.Plan(ctx, P.problem(P.CN, nx, nsteps + 1, dx, dz, 2 * np.pi / 0.1, xmin=-500.0,
                                 precision=prec, bc=P.BC_TRANSPARENT, store_dsx=1, store_dsz=1))
    db = np.zeros((1, nx, 2))
    db[0, np.abs(-500.0 + dx * np.arange(nx)) > 50.0] = (4.5e-5, 3.5e-6)  # a slab guide
    plan.material_planes(db, z_invariant=True)
    x = -500.0 + dx * np.arange(nx)
    plan.set_field(np.exp(-x ** 2 / (2 * 150.0 ** 2)) * np.exp(1j * 0.01 * x))
    plan.step(1, nsteps)
    u = plan.get_field()
    sol = plan.get_solution()
    for q in (P.Q_ABS, P.Q_ABS2, P.Q_ARG, P.Q_REAL):
        dev = plan.field_quantity(q, export_rows=5, export_cols=5)  # vectors: never resized
        assert dev.shape == (nx, 1, 1)
        _close(dev[:, 0, 0], O.field_quantity(u, q), 1e-15)
        s = plan.field_quantity(q, source=P.SRC_SOLUTION)
        assert s.shape == (nx, nsteps + 1, 1)
        _close(s[:, :, 0], O.field_quantity(sol, q), 1e-15)
        # Phase of the 2D solution is resized (src/postProcessMod.cpp:90-99)
        r = plan.field_quantity(q, source=P.SRC_SOLUTION, export_rows=40, export_cols=7)
        assert r.shape == (40, 7, 1)
        _close(r[:, :, 0], O.field_quantity(sol, q, 40, 7), 4e-15)
    plan.close()


def test_field_quantities_job_plan_and_errors(ctx):
    probs = [P.problem(P.CN, nx, 11, 1.0, 10.0, 60.0, bc=P.BC_TRANSPARENT, store_dsx=1, store_dsz=1)
             for nx in (100, 257)]
    jp = P.JobPlan(ctx, probs)
    rng = np.random.default_rng(7)
    for b, pr in enumerate(probs):
        jp.set_field(rng.standard_normal(pr.nx) + 1j * rng.standard_normal(pr.nx), b)
    jp.step(1, 10)
    for b, pr in enumerate(probs):
        dev = jp.field_quantity(P.Q_ARG, b=b)
        assert dev.shape == (pr.nx, 1, 1)
        _close(dev[:, 0, 0], O.field_quantity(jp.get_field(b), P.Q_ARG), 1e-15)
        s = jp.field_quantity(P.Q_ABS2, source=P.SRC_SOLUTION, b=b, export_rows=9, export_cols=4)
        _close(s[:, :, 0], O.field_quantity(jp.get_solution(b), P.Q_ABS2, 9, 4), 4e-15)
    jp.close()
    p = P.Plan(ctx, P.problem(P.CN, 64, 5, 1.0, 1.0, 60.0))  # stores no solution
    with pytest.raises(P.PxgError) as e:
        p.field_quantity(P.Q_ABS)  # field not set
    assert e.value.code == P.ERR_STATE
    p.set_field(np.ones(64, complex))
    assert np.array_equal(p.field_quantity(P.Q_ABS)[:, 0, 0], np.ones(64))
    with pytest.raises(P.PxgError) as e:
        p.field_quantity(P.Q_ABS, source=P.SRC_SOLUTION)
    assert e.value.code == P.ERR_STATE
    with pytest.raises(P.PxgError) as e:
        p.field_quantity(7)
    assert e.value.code == P.ERR_INVALID
    with pytest.raises(P.PxgError) as e:
        p.field_quantity(P.Q_ABS, source=3)
    assert e.value.code == P.ERR_INVALID
    # a uniform field: 255 everywhere, and the log image's 0/0 stored as 0
    assert np.all(p.field_uint8(source=P.SRC_FIELD) == 255)
    assert np.all(p.field_uint8(source=P.SRC_FIELD, log_scale=True) == 0)
    p.close()
