"""pxg_step_group: separate batch-1 FFT3D plans of one nx and precision advanced
in shared launches (one launch per pass and z-step covers every member).

Each member carries its own grid, wavenumber, material and absorber.  A
member's field must equal the same plan stepped alone by pxg_step with the
same call splits, bit for bit in both precisions (the launches run the same
kernel body per cell), and the FP64 members must match the CPU oracle to
1e-11 relative L2 (mixed precision: 1e-5)."""
import numpy as np
import pytest

import paxpro_amd as P
import pyoracle as O
from gpu_common import announce, rel_l2
from scenes import K_01, random_spheres, spheres_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = P.Ctx(0)
    yield c
    c.close()


NZ = 40


def _member(n, i):
    """Member i: its own extent, z origin, step, wavenumber, absorber and
    material (CSG phantom, vacuum, or host-sampled z-invariant planes)."""
    ext = 20000.0 * (1.0 + 0.1 * i)
    dx = ext / n
    dy = dx * (1.0 + 0.05 * (i % 2))
    dz = 10.0 + 2.0 * i
    zmin = -(NZ // 2) * dz
    k = K_01 * (1.0 + 0.07 * i)
    g = O.grid(-ext / 2, -ext / 2 + (n - 1) * dx, dx, zmin, zmin + (NZ - 1) * dz, dz, k, ymin=-ext / 2 * 0.9,
               ymax=-ext / 2 * 0.9 + (n - 1) * dy, dy=dy)
    kinds = ("csg", "csg", "none", "planes")
    kind = kinds[i % 4]
    ab = O.Absorber3D(min(4, n // 4), min(3, n // 4), 0.3, 0.2) if i % 2 == 0 else None
    x = g.xmin + g.dx * np.arange(n)
    y = g.ymin + g.dy * np.arange(n)
    w = 0.15 * ext
    u0 = np.asfortranarray(np.outer(np.exp(-((y - 0.03 * ext * i) / w) ** 2), np.exp(-(x / w) ** 2))
                           .astype(complex) * np.exp(1j * 0.3 * i))
    m = dict(g=g, kind=kind, ab=ab, u0=u0)
    if kind == "csg":
        sph = random_spheres(0x5EED0100 + i, 12, (-0.4 * ext, 0.4 * ext), (-0.3 * NZ * dz, 0.3 * NZ * dz),
                             (0.03 * ext, 0.12 * ext), (1e-6, 1e-5))
        m["csg"] = spheres_scene(sph)
    elif kind == "planes":
        X, Y = np.meshgrid(x, y)   # rows y, columns x
        d = 4e-6 * np.exp(-((X / (0.2 * ext)) ** 2 + (Y / (0.25 * ext)) ** 2))
        db = np.stack([d.ravel(order="F"), (d / 40).ravel(order="F")], axis=-1)[None]
        m["planes"] = np.ascontiguousarray(db)
    return m


def _plan(ctx, n, prec, m):
    g, ab = m["g"], m["ab"]
    absorb = (0, 0, 0.0, 0.0) if ab is None else (ab.npix_x, ab.npix_y, ab.inv_x, ab.inv_y)
    plan = P.Plan(ctx, P.problem(P.FFT3D, n, NZ, g.dx, g.dz, g.k, xmin=g.xmin, ymin=g.ymin, dy=g.dy,
                                 zmin=g.zmin, precision=prec, absorb=absorb))
    if m["kind"] == "csg":
        plan.material_csg(*m["csg"])
    elif m["kind"] == "planes":
        plan.material_planes(m["planes"], z_invariant=True)
    plan.set_field(m["u0"])
    return plan


def _oracle(m, first, nsteps):
    """The oracle's run of a CSG or vacuum member (sampled planes: None; those
    members are checked against their own pxg_step run only)."""
    if m["kind"] == "planes":
        return None
    mat = O.CSGMaterial(*m["csg"]) if m["kind"] == "csg" else O.NoMaterial()
    return O.fft3d_run(m["g"], mat, m["u0"], first, nsteps, absorber=m["ab"])


@pytest.mark.parametrize("n", [16, 64, 256, 1024])
@pytest.mark.parametrize("prec", [P.PREC_FP64, P.PREC_MIXED])
def test_group_equals_alone(ctx, n, prec):
    """Four members (two CSG phantoms, vacuum, sampled planes) advanced by two
    group calls (5 then 7 steps, the second reusing the maps of the first) equal
    each plan run alone by pxg_step with the same splits, bit for bit."""
    members = [_member(n, i) for i in range(4)]
    announce("group", n, prec)
    alone = []
    for m in members:
        plan = _plan(ctx, n, prec, m)
        plan.step(1, 5)
        plan.step(6, 7)
        alone.append(plan.get_field())
        plan.close()
    plans = [_plan(ctx, n, prec, m) for m in members]
    P.step_group(plans, 1, 5)
    P.step_group(plans, 6, 7)
    for i, (pl, ref) in enumerate(zip(plans, alone)):
        out = pl.get_field()
        announce("member", i, members[i]["kind"], "max diff", float(np.max(np.abs(out - ref))))
        np.testing.assert_array_equal(out, ref)
    if n <= 256:
        tol = 1e-11 if prec == P.PREC_FP64 else 1e-5
        for i, (pl, m) in enumerate(zip(plans, members)):
            ref = _oracle(m, 1, 12)
            if ref is None:
                continue
            err = rel_l2(pl.get_field(), ref)
            announce("member", i, "vs oracle", err)
            assert err < tol
    for pl in plans:
        pl.close()


def test_group_members_at_different_steps(ctx):
    """Members start at different z-steps (first_steps per member), a group of
    one equals pxg_step, and a member continues alone after a group call (its
    map of the last group step is reused) as if it had never been grouped."""
    n, prec = 128, P.PREC_MIXED
    members = [_member(n, i) for i in range(3)]
    starts = [1, 9, 4]
    alone = []
    for m, s0 in zip(members, starts):
        plan = _plan(ctx, n, prec, m)
        plan.step(s0, 6)
        plan.step(s0 + 6, 3)
        alone.append(plan.get_field())
        plan.close()
    plans = [_plan(ctx, n, prec, m) for m in members]
    P.step_group(plans, starts, 6)
    P.step_group([plans[0]], starts[0] + 6, 3)       # a group of one
    plans[1].step(starts[1] + 6, 3)                   # alone after the group
    P.step_group([plans[2]], [starts[2] + 6], 3)
    for pl, ref in zip(plans, alone):
        np.testing.assert_array_equal(pl.get_field(), ref)
        pl.close()


def test_group_eight_members_one_launch(ctx):
    """The largest group (8 members, every one a CSG phantom of its own)."""
    n, prec = 256, P.PREC_FP64
    members = []
    for i in range(P.MAX_GROUP):
        m = _member(n, 0 if i % 2 == 0 else 1)
        sph = random_spheres(0x5EED0200 + i, 10, (-7000, 7000), (-150, 150), (600, 2500), (1e-6, 1e-5))
        m["csg"] = spheres_scene(sph)
        m["kind"] = "csg"
        members.append(m)
    plans = [_plan(ctx, n, prec, m) for m in members]
    P.step_group(plans, 1, 10)
    for i, (pl, m) in enumerate(zip(plans, members)):
        ref = _oracle(m, 1, 10)
        err = rel_l2(pl.get_field(), ref)
        announce("member", i, "vs oracle", err)
        assert err < 1e-11
        pl.close()


def test_group_rejections(ctx):
    def plan(n=64, prec=P.PREC_FP64, batch=1, store=0, set_field=True):
        pl = P.Plan(ctx, P.problem(P.FFT3D, n, 10, 100.0, 10.0, K_01, batch=batch, precision=prec,
                                   store_dsx=1 if store else 0, store_dsz=store))
        if set_field:
            for b in range(batch):
                pl.set_field(np.ones((n, n), complex), b)
        return pl

    a, b = plan(), plan()
    cases = [
        ([a, plan(prec=P.PREC_MIXED)], P.ERR_INVALID),     # precision differs
        ([a, plan(n=128)], P.ERR_INVALID),                 # nx differs
        ([a, plan(batch=2)], P.ERR_UNSUPPORTED),           # batch 2
        ([a, plan(store=2)], P.ERR_UNSUPPORTED),           # solution storage
        ([plan(n=100)], P.ERR_UNSUPPORTED),                # not a power of two
        ([a, a], P.ERR_INVALID),                           # the same plan twice
        ([a, plan(set_field=False)], P.ERR_STATE),         # no initial field
        ([plan() for _ in range(P.MAX_GROUP + 1)], P.ERR_INVALID),
    ]
    for plans, code in cases:
        with pytest.raises(P.PxgError) as e:
            P.step_group(plans, 1, 2)
        assert e.value.code == code, e.value
    with pytest.raises(P.PxgError) as e:
        P.step_group([a, b], [1, 0], 2)
    assert e.value.code == P.ERR_INVALID
    cn = P.Plan(ctx, P.problem(P.CN, 64, 10, 10.0, 10.0, K_01))
    cn.set_field(np.ones(64, complex))
    with pytest.raises(P.PxgError) as e:
        P.step_group([a, cn], 1, 2)
    assert e.value.code == P.ERR_UNSUPPORTED
    # nothing ran: both plans still step normally
    P.step_group([a, b], 1, 2)
    np.testing.assert_array_equal(a.get_field(), b.get_field())
