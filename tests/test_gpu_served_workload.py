"""The request server on the first 40 requests of the seeded serving trace, at
the trace's own shapes: FFT3D N in {256, 512, 1024, 2048} over 20 um with
200..2000 z-steps of 10 nm and a 64-sphere phantom per request; curved-guide
CN jobs of Nx up to 8192 with 1000 steps (paxpro_amd/workloads.py,
served_trace, seed 0x5EED0004).

- Every served result equals the same job run directly by one pxg_step call:
  bit for bit in FP64; within 1e-6 relative L2 in mixed precision (each chunk
  boundary rounds the real-space field to complex64 once more).
- Served with shared launches (pxg_step_group) equals served with every chunk
  alone, bit for bit, in both precisions: groups form only from chunks of equal
  length, so every job keeps the chunk boundaries it has alone.
- Per kind and size, a prefix of the first request of that size (FFT3D: its
  first z-steps; CN: the whole job) served by the server matches the CPU oracle:
  FP64 1e-11 (FFT3D) and 1e-10 (CN); mixed 1e-5 (CN keeps complex128 state).
"""
import numpy as np
import pytest

import paxpro_amd as P
import paxpro_amd.workloads as W
import pyoracle as O
from gpu_common import announce, rel_l2

pytestmark = pytest.mark.gpu

COUNT = 40


@pytest.fixture(scope="module")
def ctx():
    c = P.Ctx(0)
    yield c
    c.close()


def _field(r):
    if r["kind"] == "fft3d":
        n = r["n"]
        return W.gaussian_field(n, 20000.0 / n, -10000.0, 3000.0)
    return np.ones(r["n"], complex)


def _serve(ctx, trace, max_group, nsteps=None, slots=8, chunk=64):
    srv = P.Server(ctx, slots=slots, chunk_steps=chunk, max_group=max_group)
    ids = [srv.submit(r["prob"], _field(r), nsteps or r["nsteps"], csg=r.get("csg"), curved=r.get("curved"))
           for r in trace]
    outs = []
    for i, jid in enumerate(ids):
        rc, _, out = srv.wait(jid, timeout=300.0)
        assert rc == P.OK, (i, rc)
        outs.append(out)
    st = srv.stats()
    srv.close()
    return outs, st


def _direct(ctx, r):
    plan = P.Plan(ctx, r["prob"])
    if "csg" in r:
        plan.material_csg(*r["csg"])
    else:
        plan.material_curved(**r["curved"])
    plan.set_field(_field(r))
    plan.step(1, r["nsteps"])
    out = plan.get_field()
    plan.close()
    return out


def test_served_trace_mixed_equals_direct_and_ungrouped(ctx):
    trace = W.served_trace(COUNT)
    sizes = sorted({(r["kind"], r["n"]) for r in trace if r["kind"] == "fft3d"})
    announce("served trace", COUNT, "fft3d sizes", sizes, "cn jobs", sum(r["kind"] == "cn" for r in trace))
    assert {n for _, n in sizes} == {256, 512, 1024, 2048}
    grouped, st = _serve(ctx, trace, max_group=P.MAX_GROUP)
    announce("grouped chunks", st["grouped_chunks"], "of", st["chunks"])
    assert st["completed"] == COUNT and st["grouped_chunks"] > 0
    alone, st1 = _serve(ctx, trace, max_group=1)
    assert st1["grouped_chunks"] == 0
    for i, r in enumerate(trace):
        np.testing.assert_array_equal(grouped[i], alone[i])
        err = rel_l2(grouped[i], _direct(ctx, r))
        assert err < 1e-6, (i, r["kind"], r["n"], r["nsteps"], err)


def test_served_trace_fp64_equals_direct(ctx):
    trace = W.served_trace(16, precision=P.PREC_FP64)
    grouped, st = _serve(ctx, trace, max_group=P.MAX_GROUP)
    announce("fp64 grouped chunks", st["grouped_chunks"], "of", st["chunks"])
    for i, r in enumerate(trace):
        np.testing.assert_array_equal(grouped[i], _direct(ctx, r), err_msg=f"request {i} {r['kind']} {r['n']}")


PREFIX = {256: 16, 512: 8, 1024: 4, 2048: 2}


@pytest.mark.parametrize("prec", [P.PREC_FP64, P.PREC_MIXED])
def test_served_prefixes_match_oracle(ctx, prec):
    trace = W.served_trace(COUNT, precision=prec)
    firsts = {}
    for r in trace:
        firsts.setdefault((r["kind"], r["n"] if r["kind"] == "fft3d" else 0), r)
    picked = [firsts[k] for k in sorted(firsts)]
    outs = []
    for r in picked:   # one server per prefix length
        steps = PREFIX[r["n"]] if r["kind"] == "fft3d" else r["nsteps"]
        out, _ = _serve(ctx, [r], max_group=P.MAX_GROUP, nsteps=steps, slots=1, chunk=max(1, steps // 2))
        outs.append((r, steps, out[0]))
    for r, steps, out in outs:
        p = r["prob"]
        g = O.grid(p.xmin, p.xmin + (p.nx - 1) * p.dx, p.dx, p.zmin, p.zmin + (p.nz - 1) * p.dz, p.dz, p.k)
        if r["kind"] == "fft3d":
            ref = O.fft3d_run(g, O.CSGMaterial(*r["csg"]), _field(r), 1, steps)
            tol = 1e-11 if prec == P.PREC_FP64 else 1e-5
        else:
            ref = O.cn_run(g, O.CurvedMaterial(**r["curved"]), _field(r), 1, steps, bc=O.BC_TRANSPARENT)
            tol = 1e-10 if prec == P.PREC_FP64 else 1e-5
        err = rel_l2(out, ref)
        announce("prefix", r["kind"], r["n"], steps, "steps, prec", prec, "rel_l2", err)
        assert err < tol
