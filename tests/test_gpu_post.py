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
