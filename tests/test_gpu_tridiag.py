"""GPU parity of the tridiagonal family (CrankNicholson, ADI, batched Thomas)
against the CPU FP64 oracle through the C ABI.

Tolerances: FP64 paths 1e-10 relative L2 (partitioned elimination and PCR
reorder the reference's serial Thomas sweep; both are FP64); the mixed CN path
(complex64 storage, FP64 sweeps) 1e-5 on field and intensity.
"""
import numpy as np
import pytest

import paxpro_amd as P
import pyoracle as O
from gpu_common import announce, rel_l2, rel_l2_int
from scenes import K_01, random_spheres, spheres_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = P.Ctx(0)
    yield c
    c.close()


def slab_planes(x, lo, hi, core=(0.0, 0.0), clad=(4.5e-5, 3.5e-6)):
    inside = (x < hi) & (x > lo)
    db = np.empty((1, x.size, 2))
    db[0, :, 0] = np.where(inside, core[0], clad[0])
    db[0, :, 1] = np.where(inside, core[1], clad[1])
    return db


@pytest.mark.parametrize("nx,nsteps,absorb", [(1024, 200, 0), (300, 100, 50), (5, 20, 0), (2, 10, 0),
                                               (3000, 20, 50), (4700, 20, 50), (9000, 10, 0)])
def test_cn_slab_vs_oracle(ctx, nx, nsteps, absorb):
    """Config-0 style slab waveguide (host-sampled material), plane wave, TBC.
    The sizes reach every march kernel: chunks of at most 4 rows held in
    registers (1024; 300: 128 chunks of 2-3 rows; 5, 2), chunks in scratch
    (3000: 512 chunks of 5-6 rows; 9000: 17-18 rows) and 9-16 rows with the
    field in registers (4700: 512 chunks of 9-10 rows)."""
    announce("cn slab", nx, nsteps, absorb)
    k, dz = K_01, 500.0
    xmin, dx = -500.0, 1000.0 / (nx - 1)
    g = O.grid(xmin, xmin + (nx - 1) * dx, dx, 0.0, nsteps * dz, dz, k)
    x = xmin + dx * np.arange(nx)
    mat = O.SlabMaterial(-50.0, 50.0, 0.0, 0.0, 4.5e-5, 3.5e-6)
    u0 = np.ones(nx, complex)
    sol_ref = np.zeros((nx, nsteps + 1), complex, order="F")
    sol_ref[:, 0] = u0
    ref = O.cn_run(g, mat, u0, 1, nsteps, bc=O.BC_TRANSPARENT, solution=sol_ref,
                   absorb_npix=absorb, absorb_inv=0.1)
    prob = P.problem(P.CN, nx, nsteps + 1, dx, dz, k, xmin=xmin, bc=P.BC_TRANSPARENT,
                     absorb=(absorb, 0, 0.1, 0.0), store_dsx=1, store_dsz=1)
    plan = P.Plan(ctx, prob)
    plan.material_planes(slab_planes(x, -50.0, 50.0), z_invariant=True)
    plan.set_field(u0)
    plan.step(1, nsteps)
    out = plan.get_field()
    err = rel_l2(out, ref)
    announce("rel_l2", err)
    assert err < 1e-10
    assert rel_l2(plan.get_solution(), sol_ref) < 1e-10


@pytest.mark.parametrize("nx,bc,absorb,material", [(1024, P.BC_TRANSPARENT, 0, "slab"),
                                                    (777, P.BC_TRANSPARENT, 50, "slab"),
                                                    (300, P.BC_DIRICHLET, 50, "slab"),
                                                    (512, P.BC_TRANSPARENT, 0, "none"),
                                                    (6, P.BC_TRANSPARENT, 0, "slab")])
def test_cn_factored_march_equals_stepwise(ctx, nx, bc, absorb, material):
    """A z-invariant CN call of many steps marches with the matrix factored once
    and the transparent-boundary rows applied by the Woodbury identity
    (k_cn_fixed); calls of one step take the plain elimination (k_cn_reg).  The
    two agree to FP64 rounding, on the exit field and on every stored step."""
    nsteps, dz = 120, 500.0
    announce("cn factored vs stepwise", nx, bc, absorb, material)
    xmin, dx = -500.0, 1000.0 / (nx - 1)
    x = xmin + dx * np.arange(nx)
    prob = P.problem(P.CN, nx, nsteps + 1, dx, dz, K_01, xmin=xmin, bc=bc,
                     absorb=(absorb, 0, 0.1, 0.0), store_dsx=1, store_dsz=1)
    u0 = np.exp(-((x + 60.0) / 150.0) ** 2) * np.exp(0.02j * x) + 0.1
    out = []
    for stepwise in (False, True):
        plan = P.Plan(ctx, prob)
        if material == "slab":
            plan.material_planes(slab_planes(x, -50.0, 50.0), z_invariant=True)
        else:
            plan.material_none()
        plan.set_field(u0)
        if stepwise:
            for st in range(1, nsteps + 1):
                plan.step(st, 1)
        else:
            plan.step(1, nsteps)
        out.append((plan.get_field(), plan.get_solution()))
        plan.close()
    err = rel_l2(out[0][0], out[1][0])
    announce("rel_l2 field", err)
    assert err < 1e-12
    assert rel_l2(out[0][1], out[1][1]) < 1e-12


def test_cn_curved_waveguide_and_dirichlet(ctx):
    nx, nsteps, dz = 777, 150, 100.0
    announce("cn curved", nx, nsteps)
    k = 2 * np.pi / 0.13
    R, w = 5e6, 120.0
    L = nsteps * dz
    lo = -L * L / (2 * R) - 2 * w - 200
    hi = w + 200
    dx = (hi - lo) / (nx - 1)
    g = O.grid(lo, hi, dx, 0.0, L, dz, k)
    kw = dict(R=R, width=w, clad_delta=4.5e-5, clad_beta=3.5e-6, dx=dx, dz=dz)
    u0 = np.ones(nx, complex)
    for bc in (O.BC_TRANSPARENT, O.BC_DIRICHLET):
        ref = O.cn_run(g, O.CurvedMaterial(**kw), u0, 1, nsteps, bc=bc)
        plan = P.Plan(ctx, P.problem(P.CN, nx, nsteps + 1, dx, dz, k, xmin=lo, bc=bc))
        plan.material_curved(**kw)
        plan.set_field(u0)
        plan.step(1, 60)
        plan.step(61, nsteps - 60)
        assert rel_l2(plan.get_field(), ref) < 1e-10


def test_cn_mixed_precision(ctx):
    nx, nsteps = 1024, 400
    announce("cn mixed", nx, nsteps)
    k, dz = K_01, 500.0
    xmin, dx = -500.0, 1000.0 / (nx - 1)
    x = xmin + dx * np.arange(nx)
    out = []
    for prec in (P.PREC_FP64, P.PREC_MIXED):
        plan = P.Plan(ctx, P.problem(P.CN, nx, nsteps + 1, dx, dz, k, xmin=xmin, precision=prec))
        plan.material_planes(slab_planes(x, -50.0, 50.0), z_invariant=True)
        plan.set_field(np.ones(nx, complex))
        plan.step(1, nsteps)
        out.append(plan.get_field())
    ef, ei = rel_l2(out[1], out[0]), rel_l2_int(out[1], out[0])
    announce("cn mixed rel_l2", ef, ei)
    assert ef < 1e-5 and ei < 1e-5


def test_cn_batched_uneven_k(ctx):
    """A batch of CN fields with different wavenumbers equals single runs."""
    nx, nsteps = 512, 50
    dz, xmin, dx = 500.0, -500.0, 1000.0 / 511
    x = xmin + dx * np.arange(nx)
    ks = np.array([K_01, 2 * np.pi / 0.05, 2 * np.pi / 0.2, 2 * np.pi / 0.12])
    plan = P.Plan(ctx, P.problem(P.CN, nx, nsteps + 1, dx, dz, K_01, xmin=xmin, batch=4))
    plan.set_wavenumbers(ks)
    plan.material_planes(slab_planes(x, -50.0, 50.0), z_invariant=True)
    for b in range(4):
        plan.set_field(np.ones(nx, complex), b)
    plan.step(1, nsteps)
    for b in range(4):
        g = O.grid(xmin, xmin + (nx - 1) * dx, dx, 0.0, nsteps * dz, dz, ks[b])
        ref = O.cn_run(g, O.SlabMaterial(-50.0, 50.0, 0.0, 0.0, 4.5e-5, 3.5e-6), np.ones(nx, complex), 1, nsteps)
        assert rel_l2(plan.get_field(b), ref) < 1e-10


@pytest.mark.parametrize("quirks", [0, P.FIX_ADI_Y_DIAG])
@pytest.mark.parametrize("n,extent,nsteps", [(64, 4000.0, 4), (300, 6000.0, 2), (1000, 20000.0, 2),
                                             (64, 128.0, 3)])
def test_adi_vs_oracle(ctx, quirks, n, extent, nsteps):
    """Dominant grids (dx = 62.5 and 20 nm: tile-Jacobi path; 300 and 1000
    cells span several tiles along the lines and a partial last group of
    lines) and a fine grid (dx = 2 nm, rho = 0.04: sequential sweep)."""
    announce("adi", n, extent, nsteps, quirks)
    k, dz = K_01, 20.0
    dx = extent / n
    xmin = -extent / 2
    g = O.grid(xmin, xmin + (n - 1) * dx, dx, -40.0, 40.0, dz, k)
    s = extent / 4000.0
    sph = random_spheres(0x5EED0002, 6, (-1500 * s, 1500 * s), (-40, 40), (300 * s, 900 * s), (1e-6, 1e-5))
    csg = spheres_scene(sph)
    xs = xmin + dx * np.arange(n)
    w = 800.0 * s
    u0 = np.asfortranarray(np.exp(-(xs[:, None] ** 2 + xs[None, :] ** 2) / w ** 2).astype(complex))
    for use_tbc in (True, False):
        ref = O.adi_run(g, O.CSGMaterial(*csg), u0, 1, nsteps, use_tbc=use_tbc, quirks=quirks)
        plan = P.Plan(ctx, P.problem(P.ADI, n, nsteps + 1, dx, dz, k, xmin=xmin, zmin=-40.0,
                                     use_tbc=int(use_tbc), quirks=quirks))
        plan.material_csg(*csg)
        plan.set_field(u0)
        plan.step(1, nsteps)
        err = rel_l2(plan.get_field(), ref)
        announce("adi rel_l2", use_tbc, err)
        assert err < 1e-10


def test_adi_planes_batch_vs_oracle(ctx):
    """Host-sampled planes (one per half-step, the CSG scene evaluated at the
    half-step's z) on a batch of two wavenumbers equal the oracle's CSG run of
    each field: checks the plane index of each half-step and the per-field
    wavenumber of the tile path."""
    announce("adi planes batch")
    n, nsteps, dz, extent = 300, 3, 20.0, 6000.0
    dx = extent / n
    xmin, zmin = -extent / 2, -40.0
    ks = np.array([K_01, 2 * np.pi / 0.15])
    sph = random_spheres(0x5EED0005, 6, (-1500, 1500), (-40, 40), (300, 900), (1e-6, 1e-5))
    csg = O.CSGMaterial(*spheres_scene(sph))
    xs = xmin + dx * np.arange(n)
    planes = np.empty((2 * nsteps, n * n, 2))
    for st in range(1, nsteps + 1):
        for half in (0, 1):
            z = zmin + st * dz + half * dz / 2
            vals = np.array([[csg.eval(xs[c], xs[r], z) for r in range(n)] for c in range(n)])
            planes[2 * (st - 1) + half] = vals.reshape(n * n, 2)
    u0 = np.asfortranarray(np.exp(-(xs[:, None] ** 2 + xs[None, :] ** 2) / 800.0 ** 2).astype(complex))
    plan = P.Plan(ctx, P.problem(P.ADI, n, nsteps + 1, dx, dz, K_01, xmin=xmin, zmin=zmin, batch=2))
    plan.set_wavenumbers(ks)
    plan.material_planes(planes, first_plane=2)
    for b in range(2):
        plan.set_field(u0, b)
    plan.step(1, nsteps)
    for b in range(2):
        g = O.grid(xmin, xmin + (n - 1) * dx, dx, zmin, zmin + nsteps * dz, dz, ks[b])
        ref = O.adi_run(g, csg, u0, 1, nsteps)
        err = rel_l2(plan.get_field(b), ref)
        announce("adi planes batch", b, err)
        assert err < 1e-10


def test_adi_mixed_precision(ctx):
    n, nsteps, extent = 256, 20, 5000.0
    k, dz = K_01, 20.0
    dx = extent / n
    xmin = -extent / 2
    csg = spheres_scene(random_spheres(0x5EED0002, 12, (-2000, 2000), (-200, 200), (300, 900), (1e-6, 1e-5)))
    xs = xmin + dx * np.arange(n)
    u0 = np.asfortranarray(np.exp(-(xs[:, None] ** 2 + xs[None, :] ** 2) / 1000.0 ** 2).astype(complex))
    out = []
    for prec in (P.PREC_FP64, P.PREC_MIXED):
        plan = P.Plan(ctx, P.problem(P.ADI, n, nsteps + 1, dx, dz, k, xmin=xmin, zmin=-200.0, precision=prec))
        plan.material_csg(*csg)
        plan.set_field(u0)
        plan.step(1, nsteps)
        out.append(plan.get_field())
    ef, ei = rel_l2(out[1], out[0]), rel_l2_int(out[1], out[0])
    announce("adi mixed", ef, ei)
    assert ef < 1e-5 and ei < 1e-5


def test_thomas_batched_matches_reference_fixtures(ctx):
    g = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                           "thomas_golden.npz"))
    for key in g["cases"]:
        d, s, r = g[f"diag_{key}"], g[f"sub_{key}"], g[f"rhs_{key}"]
        x, _ = ctx.thomas_batched(d[None, :], s[None, :], r[None, :])
        assert rel_l2(x[0], g[f"x_{key}"]) < 1e-12
    rng = np.random.default_rng(3)
    nsys, n = 300, 257
    D = 4 + rng.standard_normal((nsys, n)) + 1j * rng.standard_normal((nsys, n))
    S = rng.standard_normal((nsys, n - 1)) + 1j * rng.standard_normal((nsys, n - 1))
    B = rng.standard_normal((nsys, n)) + 1j * rng.standard_normal((nsys, n))
    xc, _ = ctx.thomas_batched(D, S, B)
    xi, _ = ctx.thomas_batched(D.T.copy(), S.T.copy(), B.T.copy(), layout=P.LAYOUT_INTERLEAVED)
    for i in (0, 17, 299):
        ref = O.thomas(D[i], S[i], B[i])
        assert rel_l2(xc[i], ref) < 1e-12
        assert rel_l2(xi[:, i], ref) < 1e-12


def test_cn_job_plan_uneven_curved_guides(ctx):
    """Config-3 style: jobs of uneven Nx, wavelengths and curved guides in one
    launch; each equals its own oracle run (FP64) / stays within 1e-5 (mixed)."""
    from scenes import splitmix64
    g = splitmix64(0x5EED0003)
    jobs = []
    for _ in range(6):
        nx = int(round(np.exp(np.log(256) + (np.log(2048) - np.log(256)) * next(g))))
        lam = 0.05 + 0.15 * next(g)
        R = 1e6 + 49e6 * next(g)
        w = 50 + 150 * next(g)
        jobs.append((nx, lam, R, w))
    nsteps, dz, L = 60, 100.0, 6000.0
    for prec, tol in ((P.PREC_FP64, 1e-10), (P.PREC_MIXED, 1e-5)):
        probs, refs = [], []
        for nx, lam, R, w in jobs:
            lo, hi = -L * L / (2 * R) - 2 * w - 200, w + 200
            dx = (hi - lo) / (nx - 1)
            probs.append(P.problem(P.CN, nx, nsteps + 1, dx, dz, 2 * np.pi / lam, xmin=lo,
                                   precision=prec, bc=P.BC_TRANSPARENT, store_dsx=1, store_dsz=1))
            gg = O.grid(lo, hi, dx, 0.0, nsteps * dz, dz, 2 * np.pi / lam)
            m = O.CurvedMaterial(R=R, width=w, clad_delta=4.5e-5, clad_beta=3.5e-6, dx=dx, dz=dz)
            refs.append(O.cn_run(gg, m, np.ones(nx, complex), 1, nsteps))
        plan = P.JobPlan(ctx, probs)
        for b, (nx, lam, R, w) in enumerate(jobs):
            plan.material_curved_job(b, R=R, width=w, clad_delta=4.5e-5, clad_beta=3.5e-6,
                                     dx=probs[b].dx, dz=dz)
            plan.set_field(np.ones(nx, complex), b)
        plan.step(1, nsteps)
        for b in range(len(jobs)):
            err = rel_l2(plan.get_field(b), refs[b])
            announce("cn job", b, jobs[b][0], prec, err)
            assert err < tol
            sol = plan.get_solution(b)
            assert sol.shape == (jobs[b][0], nsteps + 1)
            assert rel_l2(sol[:, -1], refs[b]) < tol


def test_cn_mixed_long_curved_guide_within_bound(ctx):
    """1000 steps on the curved guide of the batched-jobs workload that drifted
    to 5.7e-5 when the CN state was stored as complex64: CN plans keep
    complex128 state, so the mixed plan agrees with FP64 far below 1e-5."""
    from paxpro_amd import workloads as W
    jobs = W.curved_guide_jobs()
    out = []
    for prec in (P.PREC_FP64, P.PREC_MIXED):
        plan = P.JobPlan(ctx, [W.curved_guide_problem(j, prec) for j in jobs[:8]])
        for b, j in enumerate(jobs[:8]):
            plan.material_curved_job(b, **j["wg"])
            plan.set_field(np.ones(j["nx"], complex), b)
        plan.step(1, 1000)
        out.append([plan.get_field(b) for b in range(8)])
        plan.close()
    worst = max(rel_l2(m, r) for m, r in zip(out[1], out[0]))
    announce("cn mixed 1000 steps worst rel_l2", worst)
    assert worst < 1e-5


POLY_EQ = [[1.0, 1e-4, 1e-7, 0.0],      # F = 1 + 1e-4 x + 1e-7 z
           [1.0, 2e-4, 0.0, 0.0],       # G
           [1.0, -1e-4, 2e-8, 0.0],     # H
           [3e-6, 1e-9, 0.0, 0.0]]      # J


def _eq_planes(eqo, xmin, dx, nx, zmin, dz, planes):
    """Host samples of a ParaxialEquation where CrankNicholson evaluates it: F, G, J
    at x, H at x +- dx/2, one plane per z = zmin + p dz."""
    out = np.empty((len(planes), nx, 5))
    for q, p in enumerate(planes):
        z = zmin + p * dz
        for i in range(nx):
            x = xmin + i * dx
            F, G, _, J = eqo.values(x, z)
            out[q, i] = (F, G, eqo.values(x + 0.5 * dx, z)[2], eqo.values(x - 0.5 * dx, z)[2], J)
    return out


@pytest.mark.parametrize("nx,nsteps,bc", [(300, 40, P.BC_TRANSPARENT), (1024, 25, P.BC_TRANSPARENT),
                                          (77, 30, P.BC_DIRICHLET)])
def test_cn_custom_equation_vs_oracle(ctx, nx, nsteps, bc):
    """A ParaxialEquation with F, G, H, J varying in x and z (src/crancNicholson.cpp:41-88),
    host-sampled into coefficient planes, against the oracle evaluating the same
    equation per row; then the default equation restored."""
    announce("cn custom equation", nx, nsteps, bc)
    k, dz = K_01, 500.0
    xmin, dx = -500.0, 1000.0 / (nx - 1)
    g = O.grid(xmin, xmin + (nx - 1) * dx, dx, 0.0, 1e6, dz, k)
    x = xmin + dx * np.arange(nx)
    planes = slab_planes(x, -50.0, 50.0)
    u0 = (1.0 + 0.2 * np.exp(-(x / 120.0) ** 2)).astype(complex)
    eqo = O.PolyEq(POLY_EQ)
    ref = O.cn_run(g, O.SlabMaterial(-50, 50, 0.0, 0.0, 4.5e-5, 3.5e-6), u0, 1, nsteps, bc=bc, equation=eqo)
    plan = P.Plan(ctx, P.problem(P.CN, nx, nsteps + 1, dx, dz, k, xmin=xmin, bc=bc))
    plan.material_planes(planes, z_invariant=True)
    plan.equation_planes(_eq_planes(eqo, xmin, dx, nx, 0.0, dz, range(nsteps + 1)))
    plan.set_field(u0)
    plan.step(1, nsteps)
    err = rel_l2(plan.get_field(), ref)
    announce("rel_l2", err)
    assert err < 1e-10
    default = O.cn_run(g, O.SlabMaterial(-50, 50, 0.0, 0.0, 4.5e-5, 3.5e-6), u0, 1, nsteps, bc=bc)
    assert rel_l2(default, ref) > 1e-4   # the equation changes the result
    plan.equation_planes(None)
    plan.set_field(u0)
    plan.step(1, nsteps)
    assert rel_l2(plan.get_field(), default) < 1e-10
    with pytest.raises(P.PxgError) as e:   # planes 0..3 do not cover step 5
        plan.equation_planes(_eq_planes(eqo, xmin, dx, nx, 0.0, dz, range(4)))
        plan.step(5, 1)
    assert e.value.code == P.ERR_INVALID
    plan.close()


@pytest.mark.parametrize("prec", [P.PREC_FP64, P.PREC_MIXED])
@pytest.mark.parametrize("n", [300, 1000])
def test_adi_tile_shapes_bitwise_and_autotune(ctx, prec, n):
    """Every tile shape of both half-steps gives the same bits (each position
    sees the same K-sweep iterate, the halo covers the sweeps), with the
    transparent boundary, the reference's y-diagonal and a CSG phantom; n = 300
    and 1000 leave partial tiles at the ends of the lines and of the line
    groups.  The autotuner picks a valid shape, caches it per shape of problem,
    and its result equals the oracle's."""
    announce("adi tiles", n, prec)
    extent, nsteps, dz, k = 20.0 * n, 3, 20.0, K_01
    dx = extent / n
    xmin = -extent / 2
    s = extent / 4000.0
    csg = spheres_scene(random_spheres(0x5EED0006, 8, (-1500 * s, 1500 * s), (-40, 40), (300 * s, 900 * s),
                                       (1e-6, 1e-5)))
    xs = xmin + dx * np.arange(n)
    w = 800.0 * s
    u0 = np.asfortranarray(np.exp(-(xs[:, None] ** 2 + (0.9 * xs[None, :]) ** 2) / w ** 2).astype(complex))
    prob = P.problem(P.ADI, n, nsteps + 1, dx, dz, k, xmin=xmin, zmin=-40.0, use_tbc=1, precision=prec)
    outs = {}
    for tx in range(P.ADI_TILES):
        for ty in range(P.ADI_TILES):
            if tx != ty and tx != 0 and ty != 0:
                continue   # every shape of each half-step, paired with the other's default
            plan = P.Plan(ctx, prob)
            plan.material_csg(*csg)
            plan.set_adi_tiles(tx, ty)
            assert plan.adi_tiles() == (tx, ty)
            plan.set_field(u0)
            plan.step(1, nsteps)
            outs[(tx, ty)] = plan.get_field()
            plan.close()
    for key, out in outs.items():
        np.testing.assert_array_equal(out, outs[(0, 0)], err_msg=str(key))
    plan = P.Plan(ctx, prob)
    with pytest.raises(P.PxgError) as e:
        plan.set_adi_tiles(P.ADI_TILES, 0)
    assert e.value.code == P.ERR_INVALID
    plan.material_csg(*csg)
    picked = plan.autotune(2)
    announce("adi autotune picked", picked)
    assert all(0 <= t < P.ADI_TILES for t in picked)
    plan.set_field(u0)
    plan.step(1, nsteps)
    np.testing.assert_array_equal(plan.get_field(), outs[(0, 0)])
    plan.close()
    again = P.Plan(ctx, prob)
    again.material_csg(*csg)
    assert again.autotune(2) == picked     # cached per shape of problem
    again.close()
    g = O.grid(xmin, xmin + (n - 1) * dx, dx, -40.0, 40.0, dz, k)
    ref = O.adi_run(g, O.CSGMaterial(*csg), u0, 1, nsteps, use_tbc=True)
    err = rel_l2(outs[(0, 0)], ref)
    announce("adi tiles vs oracle", err)
    assert err < (1e-10 if prec == P.PREC_FP64 else 1e-5)
