"""Parity at the full sizes of the benchmark workloads.

Each workload runs at its real shape and step count on the GPU, in FP64 and in
mixed precision.  Mixed must stay within 1e-5 relative L2 of FP64 on the exit
field and on its intensity (the bound stated for the mixed path).  The FP64
path is compared with the CPU oracle where the oracle finishes in seconds:
over the whole run for the 2D solvers, over a prefix of the steps for the 3D
ones (the same kernels run every step).  Size-independent properties complete
the check: in vacuum the split-step propagator conserves the L2 norm, with
absorbing material the norm cannot grow.
"""
import numpy as np
import pytest

import paxpro_amd as P
import pyoracle as O
from gpu_common import announce, rel_l2, rel_l2_int
from paxpro_amd import workloads as W

pytestmark = pytest.mark.gpu

N3, STEPS3, DX3, DZ3 = 512, 1000, 20000.0 / 512, 10.0
SPHERE = [(0.0, 0.0, 0.0, 4000.0, 5e-6, 1e-7)]


@pytest.fixture(scope="module")
def ctx():
    c = P.Ctx(0)
    yield c
    c.close()


def _fft3d(ctx, prec, nsteps, csg=True):
    prob = P.problem(P.FFT3D, N3, nsteps + 1, DX3, DZ3, W.K_01, xmin=-10000.0, zmin=-5000.0,
                     precision=prec)
    plan = P.Plan(ctx, prob)
    if csg:
        plan.material_csg(*W.spheres_csg(SPHERE))
    u0 = W.gaussian_field(N3, DX3, -10000.0, 3000.0)
    plan.set_field(u0)
    plan.step(1, nsteps)
    out = plan.get_field()
    plan.close()
    return u0, out


def test_full_fft3d_sphere_512_1000_steps(ctx):
    announce("full fft3d 512^2 x 1000")
    u0, ref = _fft3d(ctx, P.PREC_FP64, STEPS3)
    _, mix = _fft3d(ctx, P.PREC_MIXED, STEPS3)
    ef, ei = rel_l2(mix, ref), rel_l2_int(mix, ref)
    announce("mixed vs fp64", ef, ei)
    assert ef < 1e-5 and ei < 1e-5
    # absorbing sphere (beta > 0): the norm can only decrease
    assert np.linalg.norm(ref) < np.linalg.norm(u0)
    # the FP64 path against the oracle over the first 20 steps (same kernels every step)
    g = O.grid(-10000.0, -10000.0 + (N3 - 1) * DX3, DX3, -5000.0, 5000.0, DZ3, W.K_01)
    oref = O.fft3d_run(g, O.CSGMaterial(*W.spheres_csg(SPHERE)), u0, 1, 20)
    _, g20 = _fft3d(ctx, P.PREC_FP64, 20)
    assert rel_l2(g20, oref) < 1e-11


def test_full_fft3d_vacuum_conserves_norm(ctx):
    announce("full fft3d vacuum norm")
    u0, ref = _fft3d(ctx, P.PREC_FP64, STEPS3, csg=False)
    _, mix = _fft3d(ctx, P.PREC_MIXED, STEPS3, csg=False)
    n0 = np.linalg.norm(u0)
    assert abs(np.linalg.norm(ref) / n0 - 1.0) < 1e-12
    assert abs(np.linalg.norm(mix) / n0 - 1.0) < 1e-6
    assert rel_l2(mix, ref) < 1e-5


def test_full_cn_slab_1024_2000_steps(ctx):
    announce("full cn 1024 x 2000")
    outs = []
    for prec in (P.PREC_FP64, P.PREC_MIXED):
        prob, db, u0 = W.cn_slab(prec, nsteps=2000)
        plan = P.Plan(ctx, prob)
        plan.material_planes(db, z_invariant=True)
        plan.set_field(u0)
        plan.step(1, 2000)
        outs.append(plan.get_field())
        plan.close()
    prob, _, u0 = W.cn_slab(P.PREC_FP64, nsteps=2000)
    g = O.grid(prob.xmin, prob.xmin + 1023 * prob.dx, prob.dx, 0.0, 2000 * prob.dz, prob.dz, prob.k)
    ref = O.cn_run(g, O.SlabMaterial(-50.0, 50.0, 0.0, 0.0, 4.5e-5, 3.5e-6), u0, 1, 2000,
                   bc=O.BC_TRANSPARENT, absorb_npix=50, absorb_inv=0.1)
    err = rel_l2(outs[0], ref)
    announce("cn fp64 vs oracle", err)
    assert err < 1e-10
    assert rel_l2(outs[1], outs[0]) < 1e-5 and rel_l2_int(outs[1], outs[0]) < 1e-5


def test_full_cn_jobs_256_x_1000_steps(ctx):
    announce("full cn jobs 256 x 1000")
    jobs = W.curved_guide_jobs()
    outs = []
    for prec in (P.PREC_FP64, P.PREC_MIXED):
        plan = P.JobPlan(ctx, [W.curved_guide_problem(j, prec) for j in jobs])
        for b, j in enumerate(jobs):
            plan.material_curved_job(b, **j["wg"])
            plan.set_field(np.ones(j["nx"], complex), b)
        plan.step(1, 1000)
        outs.append([plan.get_field(b) for b in range(len(jobs))])
        plan.close()
    worst = max(max(rel_l2(m, r), rel_l2_int(m, r)) for m, r in zip(outs[1], outs[0]))
    assert worst < 1e-5
    # four jobs (smallest, largest and two others) against the oracle
    sizes = [j["nx"] for j in jobs]
    for b in sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 0, 101}):
        j = jobs[b]
        g = O.grid(j["xmin"], j["xmin"] + (j["nx"] - 1) * j["dx"], j["dx"], 0.0, 1000 * j["dz"], j["dz"],
                   2 * np.pi / j["lam"])
        ref = O.cn_run(g, O.CurvedMaterial(**j["wg"]), np.ones(j["nx"], complex), 1, 1000)
        err = rel_l2(outs[0][b], ref)
        announce("cn job", b, j["nx"], err)
        assert err < 1e-10


def test_full_adi_2048_500_steps(ctx):
    announce("full adi 2048^2 x 500")
    csg = W.spheres_csg(W.adi_phantom())
    outs = []
    for prec in (P.PREC_FP64, P.PREC_MIXED):
        prob, u0 = W.adi_problem(prec, nsteps=500)
        plan = P.Plan(ctx, prob)
        plan.material_csg(*csg)
        plan.set_field(u0)
        plan.step(1, 500)
        outs.append(plan.get_field())
        plan.close()
    ef, ei = rel_l2(outs[1], outs[0]), rel_l2_int(outs[1], outs[0])
    announce("adi mixed vs fp64", ef, ei)
    assert ef < 1e-5 and ei < 1e-5
    # the FP64 path against the oracle over the first 2 steps
    prob, u0 = W.adi_problem(P.PREC_FP64, nsteps=500)
    g = O.grid(prob.xmin, prob.xmin + 2047 * prob.dx, prob.dx, prob.zmin, prob.zmin + 500 * prob.dz, prob.dz,
               prob.k)
    ref = O.adi_run(g, O.CSGMaterial(*csg), u0, 1, 2)
    plan = P.Plan(ctx, prob)
    plan.material_csg(*csg)
    plan.set_field(u0)
    plan.step(1, 2)
    err = rel_l2(plan.get_field(), ref)
    announce("adi fp64 vs oracle (2 steps)", err)
    assert err < 1e-10
    plan.close()
