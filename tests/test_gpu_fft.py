"""GPU parity of the split-step Fourier path (FFTSolver3D / FFTSolver2D) against
the CPU FP64 oracle, through the C ABI.

Tolerances: the FP64 path must agree with the oracle to 1e-11 relative L2 on
the exit field (both are FP64 transforms with different but equally accurate
rounding); the mixed path (complex64 fields, FP64 phase arguments) to 1e-5 on
the field and 2e-5 on the intensity, the bound stated for this path.
"""
import numpy as np
import pytest

import paxpro_amd as P
import pyoracle as O
from gpu_common import announce, fft3d_problem, rel_l2, rel_l2_int
from scenes import K_01, fft3d_case, gaussian_field, random_spheres, spheres_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = P.Ctx(0)
    yield c
    c.close()


def _run_gpu(ctx, prob, field, nsteps, csg=None, first=1, b_fields=None):
    plan = P.Plan(ctx, prob)
    if csg is not None:
        plan.material_csg(*csg)
    fields = b_fields if b_fields is not None else [field]
    for b, f in enumerate(fields):
        plan.set_field(f, b)
    plan.step(first, nsteps)
    plan.sync()
    out = [plan.get_field(b) for b in range(len(fields))]
    return plan, out


@pytest.mark.parametrize("n,nsteps", [(64, 20), (128, 12), (256, 6)])
def test_fft3d_fp64_sphere_vs_oracle(ctx, n, nsteps):
    announce("fft3d fp64 sphere", n, nsteps)
    g = fft3d_case(n, dz=10.0, nsteps=nsteps)
    u0 = gaussian_field(g, n, 3000.0)
    csg = spheres_scene([(0.0, 0.0, 0.0, 4000.0, 5e-6, 1e-7)])
    ref = O.fft3d_run(g, O.CSGMaterial(*csg), u0, 1, nsteps)
    _, (out,) = _run_gpu(ctx, fft3d_problem(g, n, nsteps + 1), u0, nsteps, csg)
    err = rel_l2(out, ref)
    announce("rel_l2", err)
    assert err < 1e-11


def test_fft3d_fp64_many_spheres_crossing_and_absorber(ctx):
    n, nsteps = 128, 10
    announce("fft3d fp64 random spheres + absorber", n, nsteps)
    g = fft3d_case(n, dz=200.0, nsteps=nsteps)
    sph = random_spheres(0x5EED0002, 24, (-8000, 8000), (-1500, 1500), (500, 3000), (1e-6, 1e-5))
    csg = spheres_scene(sph)
    ab = O.Absorber3D(6, 4, 0.3, 0.2)
    u0 = gaussian_field(g, n, 5000.0)
    ref = O.fft3d_run(g, O.CSGMaterial(*csg), u0, 1, nsteps, absorber=ab)
    _, (out,) = _run_gpu(ctx, fft3d_problem(g, n, nsteps + 1, absorber=ab), u0, nsteps, csg)
    err = rel_l2(out, ref)
    announce("rel_l2", err)
    assert err < 1e-11
    # the quirk switch changes the crossing cells
    ref_fix = O.fft3d_run(g, O.CSGMaterial(*csg), u0, 1, nsteps, absorber=ab, quirks=O.FIX_REFR_SAMPLES)
    _, (out_fix,) = _run_gpu(ctx, fft3d_problem(g, n, nsteps + 1, absorber=ab, quirks=P.FIX_REFR_SAMPLES),
                             u0, nsteps, csg)
    assert rel_l2(out_fix, ref_fix) < 1e-11
    assert rel_l2(ref_fix, ref) > 1e-9


def test_fft3d_mixed_vs_fp64(ctx):
    n, nsteps = 256, 200
    announce("fft3d mixed vs fp64", n, nsteps)
    g = fft3d_case(n, dz=10.0, nsteps=nsteps)
    u0 = gaussian_field(g, n, 3000.0)
    csg = spheres_scene([(0.0, 0.0, 0.0, 4000.0, 5e-6, 1e-7)])
    _, (ref,) = _run_gpu(ctx, fft3d_problem(g, n, nsteps + 1), u0, nsteps, csg)
    _, (mix,) = _run_gpu(ctx, fft3d_problem(g, n, nsteps + 1, precision=P.PREC_MIXED), u0, nsteps, csg)
    ef, ei = rel_l2(mix, ref), rel_l2_int(mix, ref)
    announce("mixed rel_l2 field", ef, "intensity", ei)
    assert ef < 1e-5 and ei < 2e-5


def test_fft3d_batched_equals_single(ctx):
    """Three fields with different wavenumbers in one plan equal three single runs."""
    n, nsteps = 128, 8
    announce("fft3d batch of 3", n, nsteps)
    g = fft3d_case(n, dz=10.0, nsteps=nsteps)
    csg = spheres_scene([(0.0, 0.0, 0.0, 4000.0, 5e-6, 1e-7)])
    ks = [K_01, 2 * np.pi / 0.08, 2 * np.pi / 0.15]
    fields = [gaussian_field(g, n, w) for w in (3000.0, 2500.0, 4000.0)]
    for prec, tol in ((P.PREC_FP64, 1e-13), (P.PREC_MIXED, 1e-6)):
        plan = P.Plan(ctx, fft3d_problem(g, n, nsteps + 1, precision=prec, batch=3))
        plan.set_wavenumbers(ks)
        plan.material_csg(*csg)
        for b, f in enumerate(fields):
            plan.set_field(f, b)
        plan.step(1, nsteps)
        outs = [plan.get_field(b) for b in range(3)]
        for b in range(3):
            gb = fft3d_case(n, dz=10.0, nsteps=nsteps, k=ks[b])
            _, (single,) = _run_gpu(ctx, fft3d_problem(gb, n, nsteps + 1, precision=prec), fields[b],
                                    nsteps, csg)
            assert rel_l2(outs[b], single) < tol
            if prec == P.PREC_FP64:
                ref = O.fft3d_run(gb, O.CSGMaterial(*csg), fields[b], 1, nsteps)
                assert rel_l2(outs[b], ref) < 1e-11


def test_fft3d_split_calls_and_cube_storage(ctx):
    """Two pxg_step calls equal one; the block-averaged cube matches the oracle's."""
    n, nsteps, ds, dsz = 64, 12, 8, 4
    announce("fft3d cube storage", n, nsteps)
    g = fft3d_case(n, dz=10.0, nsteps=nsteps)
    u0 = gaussian_field(g, n, 3000.0)
    csg = spheres_scene([(0.0, 0.0, 0.0, 4000.0, 5e-6, 1e-7)])
    slices = (nsteps + 1) // dsz + 1
    cube = np.zeros((n // ds, n // ds, slices), complex, order="F")
    O.cube_store(u0, cube, dsz, 0)
    ref = O.fft3d_run(g, O.CSGMaterial(*csg), u0, 1, nsteps, cube=cube, dsz=dsz)
    plan = P.Plan(ctx, fft3d_problem(g, n, nsteps + 1, store=(ds, dsz)))
    plan.material_csg(*csg)
    plan.set_field(u0)
    plan.step(1, 5)
    plan.step(6, nsteps - 5)
    assert rel_l2(plan.get_field(), ref) < 1e-11
    sol = plan.get_solution()
    assert sol.shape == cube.shape
    assert rel_l2(sol, cube) < 1e-11


def test_fft3d_vacuum_and_unsupported(ctx):
    n = 32
    g = fft3d_case(n, dz=10.0, nsteps=4)
    u0 = gaussian_field(g, n, 3000.0)
    ref = O.fft3d_run(g, O.NoMaterial, u0, 1, 4)
    _, (out,) = _run_gpu(ctx, fft3d_problem(g, n, 5), u0, 4)
    assert rel_l2(out, ref) < 1e-12
    with pytest.raises(P.PxgError) as e:
        P.Plan(ctx, P.problem(P.FFT3D, 4100, 5, 1.0, 1.0, 1.0))   # beyond 4096 per side
    assert e.value.code == P.ERR_UNSUPPORTED
    plan = P.Plan(ctx, fft3d_problem(g, n, 5))
    with pytest.raises(P.PxgError) as e:
        plan.step(1, 1)
    assert e.value.code == P.ERR_STATE


@pytest.mark.parametrize("material", ["curved", "csg"])
def test_fft2d_vs_oracle(ctx, material):
    nx, nsteps = 1024, 50
    announce("fft2d", material, nx, nsteps)
    k = K_01
    xmin, dx, dz = -600.0, 800.0 / (nx - 1), 100.0
    g = O.grid(xmin, xmin + (nx - 1) * dx, dx, 0.0, nsteps * dz, dz, k)
    x = xmin + dx * np.arange(nx)
    u0 = np.exp(-((x - 50) / 80.0) ** 2).astype(complex)
    prob = P.problem(P.FFT2D, nx, nsteps + 1, dx, dz, k, xmin=xmin, store_dsx=1, store_dsz=1)
    plan = P.Plan(ctx, prob)
    if material == "curved":
        kw = dict(R=2e6, width=100.0, clad_delta=4.5e-5, clad_beta=3.5e-6, dx=dx, dz=dz)
        m = O.CurvedMaterial(**kw)
        plan.material_curved(**kw)
    else:
        csg = spheres_scene([(0.0, 0.0, 2000.0, 150.0, 4e-6, 1e-7), (-200.0, 0.0, 3000.0, 100.0, 8e-6, 3e-7)])
        m = O.CSGMaterial(*csg)
        plan.material_csg(*csg)
    sol_ref = np.zeros((nx, nsteps + 1), complex, order="F")
    sol_ref[:, 0] = u0
    ref = O.fft2d_run(g, m, u0, 1, nsteps, solution=sol_ref)
    plan.set_field(u0)
    plan.step(1, nsteps)
    assert rel_l2(plan.get_field(), ref) < 1e-11
    assert rel_l2(plan.get_solution(), sol_ref) < 1e-11


@pytest.mark.parametrize("prec", [P.PREC_FP64, P.PREC_MIXED])
def test_fft3d_work_group_shapes_bitwise_and_autotune(ctx, prec):
    """Every supported lines-per-work-group shape of both passes gives the same
    bits (a line's arithmetic does not depend on its work-group; CSG culling is
    conservative), and the autotuner picks a supported shape, caches it in the
    context and zeroes the field it timed on."""
    n, nsteps = 128, 6
    announce("fft3d work-group shapes", n, nsteps, "precision", prec)
    g = fft3d_case(n, dz=200.0, nsteps=nsteps)
    sph = random_spheres(0x5EED0003, 12, (-8000, 8000), (-1500, 1500), (500, 3000), (1e-6, 1e-5))
    csg = spheres_scene(sph)
    u0 = gaussian_field(g, n, 5000.0)
    prob = fft3d_problem(g, n, nsteps + 1, precision=prec, batch=2)
    outs = {}
    for lines in (1, 2, 4, 8, 16):
        plan = P.Plan(ctx, prob)
        plan.material_csg(*csg)
        plan.set_lines(lines, lines)
        assert plan.lines() == (lines, lines)
        plan.set_field(u0, 0)
        plan.set_field(u0, 1)
        plan.step(1, nsteps)
        outs[lines] = plan.get_field(1)
        plan.close()
    for lines, out in outs.items():
        assert np.array_equal(out, outs[1]), lines
    plan = P.Plan(ctx, prob)
    plan.material_csg(*csg)
    with pytest.raises(P.PxgError) as e:
        plan.set_lines(32, 0)
    assert e.value.code == P.ERR_INVALID
    lx, ly = plan.autotune(2)
    assert lx in outs and ly in outs
    with pytest.raises(P.PxgError) as e:  # the field was zeroed and must be set again
        plan.step(1, 1)
    assert e.value.code == P.ERR_STATE
    plan.set_field(u0, 0)
    plan.set_field(u0, 1)
    plan.step(1, nsteps)
    assert np.array_equal(plan.get_field(1), outs[1])
    plan2 = P.Plan(ctx, prob)
    plan2.material_csg(*csg)
    assert plan2.autotune(2) == (lx, ly)  # cached for the same shape of problem
    plan.close()
    plan2.close()


def test_fft3d_autotune_batch_halves_and_plane_range(ctx):
    """The autotuner also times a batch of 16 fields marched in halves on two
    streams against one launch per pass and keeps the faster; either way the
    fields are the same bits as an untuned plan's.  For host-sampled material
    planes that change with z it times at the first steps the planes cover
    (here steps 5..7 of a 41-step plan, not the middle of the z range), and the
    march afterwards equals an untuned plan's."""
    n, nsteps = 64, 5
    announce("fft3d autotune halves and plane range", n)
    g = fft3d_case(n, dz=200.0, nsteps=40)
    sph = random_spheres(0x5EED0007, 6, (-6000, 6000), (-1000, 1000), (800, 2500), (1e-6, 1e-5))
    csg = spheres_scene(sph)
    u0 = gaussian_field(g, n, 5000.0)
    prob = fft3d_problem(g, n, 41, precision=P.PREC_MIXED, batch=16)
    outs = []
    for tune in (False, True):
        plan = P.Plan(ctx, prob)
        plan.material_csg(*csg)
        if tune:
            plan.autotune(2)
        for b in range(16):
            plan.set_field(u0 * (1.0 + 0.01 * b), b)
        plan.step(1, nsteps)
        outs.append([plan.get_field(b) for b in (0, 9, 15)])
        plan.close()
    for x, y in zip(*outs):
        assert np.array_equal(x, y)

    rng = np.random.default_rng(7)
    db = np.zeros((3, n * n, 2))
    db[:, :, 0] = 4e-6 * rng.random((3, n * n))
    db[:, :, 1] = 1e-7 * rng.random((3, n * n))
    prob1 = fft3d_problem(g, n, 41, precision=P.PREC_FP64)
    fields = []
    for tune in (False, True):
        plan = P.Plan(ctx, prob1)
        plan.material_planes(db, first_plane=5)
        if tune:
            plan.autotune(2)
        plan.set_field(u0)
        plan.step(5, 3)
        fields.append(plan.get_field())
        plan.close()
    assert np.array_equal(fields[0], fields[1])


def test_context_outlived_by_its_plans():
    """A context destroyed while a plan of it lives stays usable by that plan
    and is freed with it; no new plan can be made from it."""
    n = 32
    g = fft3d_case(n, dz=10.0, nsteps=4)
    u0 = gaussian_field(g, n, 3000.0)
    c = P.Ctx(0)
    plan = P.Plan(c, fft3d_problem(g, n, 5))
    h = c.h
    c.close()
    plan.set_field(u0)
    plan.step(1, 4)
    assert rel_l2(plan.get_field(), O.fft3d_run(g, O.NoMaterial, u0, 1, 4)) < 1e-12
    plan.close()
    assert h is not None


@pytest.mark.parametrize("prec", [P.PREC_FP64, P.PREC_MIXED])
def test_fft3d_batch_map_built_in_pass_x(ctx, prec):
    """From 8 fields on the material map is built inside pass X instead of its
    own launch: a batch of 8 over 24 spheres with border crossings gives, field
    by field, the same bits as single-field plans (whose map has its own
    launch), and the FP64 result matches the oracle."""
    n, nsteps = 128, 10
    announce("fft3d batch 8, map in pass X", n, nsteps, "precision", prec)
    g = fft3d_case(n, dz=200.0, nsteps=nsteps)
    csg = spheres_scene(random_spheres(0x5EED0002, 24, (-8000, 8000), (-1500, 1500), (500, 3000), (1e-6, 1e-5)))
    ks = [K_01 * (1 + 0.01 * b) for b in range(8)]
    u0 = gaussian_field(g, n, 5000.0)
    plan = P.Plan(ctx, fft3d_problem(g, n, nsteps + 1, precision=prec, batch=8))
    plan.set_wavenumbers(ks)
    plan.material_csg(*csg)
    for b in range(8):
        plan.set_field(u0, b)
    plan.step(1, nsteps)
    for b in (0, 5):
        gb = fft3d_case(n, dz=200.0, nsteps=nsteps, k=ks[b])
        _, (single,) = _run_gpu(ctx, fft3d_problem(gb, n, nsteps + 1, precision=prec), u0, nsteps, csg)
        assert np.array_equal(plan.get_field(b), single), b
        if prec == P.PREC_FP64:
            ref = O.fft3d_run(gb, O.CSGMaterial(*csg), u0, 1, nsteps)
            assert rel_l2(single, ref) < 1e-11
    plan.close()


@pytest.mark.parametrize("n,nsteps", [(16, 6), (4096, 2)])
def test_fft3d_smallest_and_largest_grid(ctx, n, nsteps):
    """The smallest (16^2) and largest (4096^2) supported grids: FP64 against
    the oracle, mixed against FP64."""
    announce("fft3d edge size", n, nsteps)
    g = fft3d_case(n, dz=10.0, nsteps=nsteps)
    u0 = gaussian_field(g, n, 3000.0)
    csg = spheres_scene([(0.0, 0.0, 0.0, 4000.0, 5e-6, 1e-7), (2500.0, -1500.0, 0.0, 1200.0, 8e-6, 2e-7)])
    ref = O.fft3d_run(g, O.CSGMaterial(*csg), u0, 1, nsteps)
    _, (out,) = _run_gpu(ctx, fft3d_problem(g, n, nsteps + 1), u0, nsteps, csg)
    err = rel_l2(out, ref)
    announce("rel_l2", err)
    assert err < 1e-11
    _, (mix,) = _run_gpu(ctx, fft3d_problem(g, n, nsteps + 1, precision=P.PREC_MIXED), u0, nsteps, csg)
    assert rel_l2(mix, out) < 1e-5


@pytest.mark.parametrize("batch,prec", [(1, P.PREC_FP64), (8, P.PREC_MIXED)])
def test_fft3d_graph_replays_equal_eager(ctx, batch, prec):
    """pxg_plan_set_graphs: every replay of a captured march equals the eager
    enqueue bit for bit, with eager work between replays; a material change forces
    a new capture (the key holds the kernels' arguments), and cube storage and the
    pass-X material map (batches of 8) are inside the graph."""
    announce("fft3d graphs", batch, prec)
    n, nsteps = 64, 12
    g = fft3d_case(n, dz=10.0, nsteps=nsteps)
    u0 = gaussian_field(g, n, 3000.0)
    csg1 = spheres_scene([(0.0, 0.0, 0.0, 4000.0, 5e-6, 1e-7)])
    csg2 = spheres_scene([(500.0, 0.0, 0.0, 3000.0, 8e-6, 2e-7)])
    prob = fft3d_problem(g, n, nsteps + 1, precision=prec, batch=batch, store=(4, 3))

    def run(plan, first, count):
        for b in range(batch):
            plan.set_field(u0, b)
        plan.step(first, count)
        plan.sync()
        return [plan.get_field(b) for b in range(batch)], plan.get_solution(0)

    eager = P.Plan(ctx, prob)
    graph = P.Plan(ctx, prob)
    graph.set_graphs(True)
    for csg in (csg1, csg2):
        eager.material_csg(*csg)
        graph.material_csg(*csg)
        ref = run(eager, 1, nsteps)
        for rep in range(3):
            out = run(graph, 1, nsteps)
            assert all(np.array_equal(a, r) for a, r in zip(out[0], ref[0])), rep
            assert np.array_equal(out[1], ref[1]), rep
            run(eager, 1, 5)  # eager work on the device between replays
        # a call of another shape is captured on its own
        a5 = run(graph, 1, 5)
        e5 = run(eager, 1, 5)
        assert all(np.array_equal(a, r) for a, r in zip(a5[0], e5[0]))
    graph.set_graphs(False)
    out = run(graph, 1, nsteps)
    assert all(np.array_equal(a, r) for a, r in zip(out[0], ref[0]))
    eager.close()
    graph.close()


@pytest.mark.parametrize("prec,batch", [(P.PREC_FP64, 16), (P.PREC_MIXED, 16), (P.PREC_MIXED, 64)])
def test_fft3d_batch_halves_on_two_streams(ctx, prec, batch):
    """Batches of 16 fields or more march as two halves on two streams, each half
    with its own material map; a profiled call marches the whole batch in one
    launch per pass.  Calls that alternate between the two (steps 1-4 halves,
    5-7 one launch, 8-12 halves) give, field by field, the same bits as the same
    three calls with one launch per pass (the calls' boundaries stay the same:
    each call ends with the field in real space, which rounds complex64 fields),
    and the FP64 fields match the oracle."""
    n, nsteps = 128, 12
    announce("fft3d batch halves", n, nsteps, "precision", prec, "batch", batch)
    g = fft3d_case(n, dz=200.0, nsteps=nsteps)
    csg = spheres_scene(random_spheres(0x5EED0002, 24, (-8000, 8000), (-1500, 1500), (500, 3000), (1e-6, 1e-5)))
    ks = [K_01 * (1 + 0.01 * b) for b in range(batch)]
    u0 = gaussian_field(g, n, 5000.0)
    runs = []
    for mode in ("alternating", "one launch"):
        plan = P.Plan(ctx, fft3d_problem(g, n, nsteps + 1, precision=prec, batch=batch))
        plan.set_wavenumbers(ks)
        plan.material_csg(*csg)
        for b in range(batch):
            plan.set_field(u0, b)
        if mode == "alternating":
            plan.step(1, 4)
            plan.step_profiled(5, 3)
            plan.step(8, 5)
        else:
            plan.step_profiled(1, 4)
            plan.step_profiled(5, 3)
            plan.step_profiled(8, 5)
        plan.sync()
        runs.append([plan.get_field(b) for b in range(batch)])
        plan.close()
    for b in range(batch):
        assert np.array_equal(runs[0][b], runs[1][b]), b
    if prec == P.PREC_FP64:
        for b in (0, batch // 2, batch - 1):
            gb = fft3d_case(n, dz=200.0, nsteps=nsteps, k=ks[b])
            ref = O.fft3d_run(gb, O.CSGMaterial(*csg), u0, 1, nsteps)
            assert rel_l2(runs[0][b], ref) < 1e-11, b


def test_fft3d_map_ring_across_call_splits(ctx):
    """Plans of fewer than 8 fields build the material maps of up to MAP_RING
    consecutive steps in one launch (the parts culled once, each step's z0
    plane the previous step's z1 plane).  Calls of 3, 9, 8 and 1 steps (groups
    cut by the calls, the previous map reused or rebuilt at each call) over 24
    spheres with border crossings match the oracle's single march."""
    n, dz = 128, 200.0
    calls = [(1, 3), (4, 9), (13, 8), (21, 1)]
    nsteps = sum(k for _, k in calls)
    announce("fft3d map ring across calls", n, nsteps)
    g = fft3d_case(n, dz=dz, nsteps=nsteps)
    csg = spheres_scene(random_spheres(0x5EED0002, 24, (-8000, 8000), (-1500, 1500), (500, 3000), (1e-6, 1e-5)))
    u0 = gaussian_field(g, n, 5000.0)
    plan = P.Plan(ctx, fft3d_problem(g, n, nsteps + 1))
    plan.material_csg(*csg)
    plan.set_field(u0)
    for first, k in calls:
        plan.step(first, k)
    out = plan.get_field()
    plan.close()
    ref = O.fft3d_run(g, O.CSGMaterial(*csg), u0, 1, nsteps)
    assert rel_l2(out, ref) < 1e-11
