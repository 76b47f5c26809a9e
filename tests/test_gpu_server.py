"""Serving loop (pxg_server): mixed kinds and sizes admitted by HBM budget run
concurrently; each result equals the same job run alone.  Failure handling is
forced from the host: budget rejection, invalid job, cancellation and an
injected failure, while the other jobs still complete identically."""
import numpy as np
import pytest

import paxpro_amd as P
import pyoracle as O
from gpu_common import announce, fft3d_problem
from scenes import K_01, fft3d_case, gaussian_field, random_spheres, spheres_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = P.Ctx(0)
    yield c
    c.close()


def _jobs():
    jobs = []
    for i, (n, prec, seed) in enumerate([(64, P.PREC_FP64, 1), (128, P.PREC_MIXED, 2), (128, P.PREC_FP64, 3),
                                         (256, P.PREC_MIXED, 4)]):
        g = fft3d_case(n, dz=10.0, nsteps=24, k=K_01 * (1 + 0.01 * i))
        sph = random_spheres(0x5EED0004 + seed, 16, (-8000, 8000), (-100, 100), (500, 3000), (1e-6, 1e-5))
        prob = fft3d_problem(g, n, 25, precision=prec)
        jobs.append(dict(prob=prob, field=gaussian_field(g, n, 3000.0, k=g.k), nsteps=24,
                         csg=spheres_scene(sph)))
    for i, nx in enumerate([300, 1111]):
        R, w, dz, nsteps = 5e6 + 1e6 * i, 100.0 + 20 * i, 100.0, 40
        L = nsteps * dz
        lo, hi = -L * L / (2 * R) - 2 * w - 200, w + 200
        dx = (hi - lo) / (nx - 1)
        prob = P.problem(P.CN, nx, nsteps + 1, dx, dz, 2 * np.pi / (0.08 + 0.02 * i), xmin=lo,
                         precision=P.PREC_MIXED if i else P.PREC_FP64)
        jobs.append(dict(prob=prob, field=np.ones(nx, complex), nsteps=nsteps,
                         curved=dict(R=R, width=w, clad_delta=4.5e-5, clad_beta=3.5e-6, dx=dx, dz=dz)))
    n = 64
    g = fft3d_case(n, dz=20.0, nsteps=6)
    prob = P.problem(P.ADI, n, 7, g.dx, g.dz, g.k, xmin=g.xmin, zmin=g.zmin)
    jobs.append(dict(prob=prob, field=gaussian_field(g, n, 3000.0), nsteps=6,
                     csg=spheres_scene([(0.0, 0.0, 0.0, 4000.0, 5e-6, 1e-7)])))
    return jobs


def _alone(ctx, j):
    plan = P.Plan(ctx, j["prob"])
    if "csg" in j:
        plan.material_csg(*j["csg"])
    elif "curved" in j:
        plan.material_curved(**j["curved"])
    plan.set_field(j["field"])
    plan.step(1, j["nsteps"])
    out = plan.get_field()
    plan.close()
    return out


def _served_alone(ctx, j, chunk):
    srv = P.Server(ctx, slots=1, chunk_steps=chunk)
    jid = srv.submit(j["prob"], j["field"], j["nsteps"], csg=j.get("csg"), curved=j.get("curved"))
    rc, _, out = srv.wait(jid, timeout=120.0)
    srv.close()
    assert rc == P.OK
    return out


def test_server_mix_equals_alone(ctx):
    """Each job's result in the mix is bit-identical to the same job served
    alone, and agrees with a direct single-call run: bit-identical in FP64; in
    mixed precision the chunk boundaries round the real-space field to
    complex64 once more, so the agreement is within the 1e-6 stated for it."""
    jobs = _jobs()
    chunk = 4
    direct = [_alone(ctx, j) for j in jobs]
    alone = [_served_alone(ctx, j, chunk) for j in jobs]
    srv = P.Server(ctx, slots=3, chunk_steps=chunk)
    ids = [srv.submit(j["prob"], j["field"], j["nsteps"], csg=j.get("csg"), curved=j.get("curved"))
           for j in jobs]
    for i, jid in enumerate(ids):
        rc, lat, out = srv.wait(jid, timeout=120.0)
        announce("server job", i, "rc", rc, "latency ms", round(lat, 2))
        assert rc == P.OK
        np.testing.assert_array_equal(out, alone[i])
        if jobs[i]["prob"].precision == P.PREC_FP64:
            np.testing.assert_array_equal(out, direct[i])
        else:
            assert np.linalg.norm(out - direct[i]) / np.linalg.norm(direct[i]) < 1e-6
    st = srv.stats()
    assert st["completed"] == len(jobs) and st["failed"] == 0
    assert 0 < st["p50_ms"] <= st["p99_ms"]
    srv.close()


def test_server_failures_are_isolated(ctx):
    jobs = _jobs()[:3]
    alone = [_served_alone(ctx, j, 2) for j in jobs]
    small_budget = 200 << 20
    srv = P.Server(ctx, slots=2, budget=small_budget, chunk_steps=2)
    big = P.problem(P.FFT3D, 4096, 3, 1.0, 1.0, K_01, batch=1)   # 4096^2 c128 > budget
    u_big = np.zeros((4096, 4096), complex, order="F")
    bad = P.problem(P.FFT3D, 4100, 3, 1.0, 1.0, K_01)             # beyond 4096 points per side
    ids = {}
    ids["ok0"] = srv.submit(jobs[0]["prob"], jobs[0]["field"], jobs[0]["nsteps"], csg=jobs[0]["csg"])
    ids["big"] = srv.submit(big, u_big, 2)
    ids["bad"] = srv.submit(bad, np.zeros((2, 2), complex), 2)   # rejected before its field is read
    ids["inj"] = srv.submit(jobs[1]["prob"], jobs[1]["field"], jobs[1]["nsteps"], csg=jobs[1]["csg"], inject=1)
    ids["cancel"] = srv.submit(jobs[2]["prob"], jobs[2]["field"], 2000, csg=jobs[2]["csg"])
    ids["ok2"] = srv.submit(jobs[2]["prob"], jobs[2]["field"], jobs[2]["nsteps"], csg=jobs[2]["csg"])
    srv.cancel(ids["cancel"])
    res = {k: srv.wait(v, timeout=120.0) for k, v in ids.items()}
    announce("failure codes", {k: r[0] for k, r in res.items()})
    assert res["big"][0] == P.ERR_BUDGET
    assert res["bad"][0] == P.ERR_UNSUPPORTED
    assert res["inj"][0] == P.ERR_INJECTED
    assert res["cancel"][0] == P.ERR_CANCELLED
    np.testing.assert_array_equal(res["ok0"][2], alone[0])
    np.testing.assert_array_equal(res["ok2"][2], alone[2])
    st = srv.stats()
    assert st["rejected"] == 1 and st["cancelled"] == 1 and st["completed"] == 2
    srv.close()


def test_server_plan_failures_release_memory(ctx):
    """A job whose plan cannot be recycled (pxg_plan_reset failure forced from
    the host) or created (creation failure forced from the host) fails with
    that code; the pooled plan of the failed recycle is destroyed and its bytes
    released; later jobs of both shapes finish bit-identically to the same jobs
    served alone; the context holds no bytes once the server is closed."""
    jobs = _jobs()[:2]   # 64^2 FP64 and 128^2 mixed FFT3D with CSG phantoms
    alone = [_served_alone(ctx, j, 4) for j in jobs]
    base = ctx.reserved()
    srv = P.Server(ctx, slots=1, chunk_steps=4)

    def run(j, **kw):
        jid = srv.submit(j["prob"], j["field"], j["nsteps"], csg=j["csg"], **kw)
        return srv.wait(jid, timeout=120.0)

    rc, _, out = run(jobs[0])
    assert rc == P.OK
    np.testing.assert_array_equal(out, alone[0])
    assert srv.stats()["hbm_reserved"] > 0            # its plan waits in the pool
    rc, _, _ = run(jobs[0], inject_plan=P.INJECT_RESET)
    announce("forced recycle failure", rc, srv.stats()["hbm_reserved"])
    assert rc == P.ERR_INJECTED
    assert srv.stats()["hbm_reserved"] == 0           # destroyed, bytes released
    rc, _, _ = run(jobs[1], inject_plan=P.INJECT_CREATE)
    assert rc == P.ERR_INJECTED
    assert srv.stats()["hbm_reserved"] == 0
    for j, ref in zip(jobs, alone):
        rc, _, out = run(j)
        assert rc == P.OK
        np.testing.assert_array_equal(out, ref)
    st = srv.stats()
    assert st["failed"] == 2 and st["completed"] == 3
    srv.close()
    assert ctx.reserved() == base
