#!/usr/bin/env python3
"""Benchmark of the PaxPro propagation hot path on one MI355X.

Workload: 3D split-step Fourier propagation (FFTSolver3D) on a 512 x 512
transverse grid over 20 x 20 um (dx = 39.0625 nm), 1000 z-steps of 10 nm over
z in [-5, 5] um, lambda = 0.1 nm, Gaussian beam of waist 3 um, a sphere of
radius 4 um with delta = 5e-6, beta = 1e-7 in vacuum (device-evaluated CSG).

A "step" is one z-step of every field in the batch.  The headline value is
the batched mixed-precision path (complex64 fields, FP64 phase arguments):
B fields (photon energies lambda = 0.1 nm * (1 + 0.002 b)) advance together.
Beside it the plain FP64 single-field path is timed on the same workload and
the mixed result of field 0 (lambda = 0.1 nm) is compared with the FP64 one
(relative L2 on field and intensity).  Inputs are resident in HBM before the
timed region starts.

Prints ONE JSON line on stdout; diagnostics go to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import paxpro_amd as P  # noqa: E402

N = 512
DX = 20000.0 / 512
XMIN = -10000.0
DZ = 10.0
ZMIN = -5000.0
LAMBDA = 0.1
WAIST = 3000.0
SPHERE = (0.0, 0.0, 0.0, 4000.0, 5e-6, 1e-7)
HBM_PEAK_GBS = 8000.0
# rocprofv3 counter summary of this workload: HBM bytes per launch of each pass
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "fft3d_passes_512x512x64_c64_traffic.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gaussian_field(n):
    """GaussianBeam::get at z = z0 = 0 with centre 0 and no tilt reduces to
    exp(-(x/w)^2) exp(-(y/w)^2) (src/gaussianBeam.cpp:34-63); values(j, i) =
    src(x_i, y_j) with rows = y (src/paraxialSimulation.cpp:381-395)."""
    x = XMIN + DX * np.arange(n)
    gx = np.exp(-(x / WAIST) ** 2)
    return np.asfortranarray(np.outer(gx, gx).astype(np.complex128))


def sphere_csg():
    shapes = np.zeros(1, P.SHAPE_DTYPE)
    T = np.eye(4)
    T[0, 3], T[1, 3], T[2, 3] = -SPHERE[0], -SPHERE[1], -SPHERE[2]
    shapes[0]["type"], shapes[0]["op"], shapes[0]["T"], shapes[0]["p"] = 0, 0, T.ravel(), (SPHERE[3], 0, 0)
    parts = np.zeros(1, P.PART_DTYPE)
    parts[0]["first_shape"], parts[0]["n_shapes"], parts[0]["op"] = 0, 1, 0
    parts[0]["delta"], parts[0]["beta"] = SPHERE[4], SPHERE[5]
    modules = np.zeros(1, P.MODULE_DTYPE)
    modules[0]["first_part"], modules[0]["n_parts"] = 0, 1
    return shapes, parts, modules


def make_plan(ctx, precision, batch, nz):
    k = 2 * np.pi / LAMBDA
    prob = P.problem(P.FFT3D, N, nz, DX, DZ, k, xmin=XMIN, zmin=ZMIN, precision=precision,
                     batch=batch)
    plan = P.Plan(ctx, prob)
    plan.set_wavenumbers([2 * np.pi / (LAMBDA * (1 + 0.002 * b)) for b in range(batch)])
    plan.material_csg(*sphere_csg())
    return plan


def load_fields(plan, u0, batch):
    for b in range(batch):
        plan.set_field(u0, b)


def timed(plan, steps):
    plan.sync()
    t0 = time.perf_counter()
    plan.step(1, steps)
    plan.sync()
    return time.perf_counter() - t0


def copy_calibration(mib, iters=20):
    """A plain device-to-device copy of known size; hardware counters of this
    kernel calibrate FETCH_SIZE/WRITE_SIZE against known bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(iters):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / iters
    print(json.dumps({"copy_bytes": n * 4, "iters": iters + 1, "mean_us": round(ms * 1e3, 2),
                      "read_plus_write_GBs": round(2 * n * 4 / (ms * 1e-3) / 1e9, 1)}), flush=True)


def measured_traffic(kernel, batch):
    """HBM bytes per launch of `kernel` (the plan's record name) on this
    workload, from the committed rocprofv3 summary: FETCH_SIZE and WRITE_SIZE
    per dispatch, each scaled by that counter's reading for a plain copy of
    known size.  None when no summary covers this kernel and batch."""
    try:
        with open(TRAFFIC_SUMMARY) as fh:
            lines = fh.read().splitlines()
    except OSError:
        return None
    for line in lines:
        if line.startswith("{"):
            d = json.loads(line)
            if d.get("kernel") == kernel and d.get("batch") == batch and d.get("grid") == N:
                return d["traffic_bytes_per_launch"]
    return None


def cpu_baseline(target_s):
    """The oracle (FP64 C restatement, OpenMP over cells, single-threaded FFT
    as in the reference) on the same 512^2 sphere workload for a bounded number
    of z-steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    k = 2 * np.pi / LAMBDA
    g = O.grid(XMIN, XMIN + (N - 1) * DX, DX, ZMIN, ZMIN + 1000 * DZ, DZ, k)
    shapes, parts, modules = sphere_csg()
    mat = O.CSGMaterial(shapes, parts, modules)
    u0 = gaussian_field(N)
    t0 = time.perf_counter()
    O.fft3d_run(g, mat, u0, 1, 1)
    one = time.perf_counter() - t0
    steps = max(1, min(1000, int(target_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    O.fft3d_run(g, mat, u0, 1, steps)
    dt = time.perf_counter() - t0
    return {"value": N * N * steps / dt / 1e6, "unit": "Mcell*steps/s", "cores": O.num_threads(),
            "kind": "port",
            "sample": f"oracle FFTSolver3D restatement, 512x512 sphere workload, {steps} z-steps "
                      f"(of 1000) in {dt:.1f} s, FP64, self-written radix-2 FFT (not FFTW)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="z-steps of the HIP-event profile (default: all --steps)")
    ap.add_argument("--mixed-only", action="store_true",
                    help="time only the batched mixed path (profiling runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run a plain device copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(sharding is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    nz = max(args.steps, args.warmup) + 1
    if args.profile_steps <= 0:
        args.profile_steps = args.steps
    B = args.batch
    ctx = P.Ctx(0)
    u0 = gaussian_field(N)

    # --- plain FP64 single-solve path
    if not args.mixed_only:
        p64 = make_plan(ctx, P.PREC_FP64, 1, nz)
        load_fields(p64, u0, 1)
        if args.warmup:
            timed(p64, args.warmup)
            load_fields(p64, u0, 1)
        t64 = timed(p64, args.steps)
        ref = p64.get_field(0)
        v64 = N * N * args.steps / t64 / 1e6
        log(f"fp64 single: {t64 * 1e3:.2f} ms for {args.steps} steps -> {v64:.1f} Mcell*steps/s")
        p64.close()

    # --- batched mixed-precision path (headline)
    pm = make_plan(ctx, P.PREC_MIXED, B, nz)
    load_fields(pm, u0, B)
    if args.warmup:
        timed(pm, args.warmup)
        load_fields(pm, u0, B)
    tm = timed(pm, args.steps)
    mix = pm.get_field(0)
    vm = B * N * N * args.steps / tm / 1e6
    log(f"mixed batch {B}: {tm * 1e3:.2f} ms -> {vm:.1f} Mcell*steps/s")
    if args.mixed_only:
        load_fields(pm, u0, B)
        _, stats = pm.step_profiled(1, args.profile_steps)
        for name, s in stats.items():
            log(f"  {name}: {s['launches']} launches, mean {s['mean_ms'] * 1e3:.1f} us, "
                f"{s['bytes_per_launch'] / s['mean_ms'] / 1e6:.1f} GB/s alg")
        pm.close()
        print(json.dumps({"mixed_value": vm, "batch": B, "steps": args.steps}), flush=True)
        return
    ef = float(np.linalg.norm(mix - ref) / np.linalg.norm(ref))
    ir, im = np.abs(ref) ** 2, np.abs(mix) ** 2
    ei = float(np.linalg.norm(im - ir) / np.linalg.norm(ir))
    log(f"  rel L2 vs fp64: field {ef:.3e} intensity {ei:.3e}")

    # --- per-kernel durations with HIP events on the plan's stream
    load_fields(pm, u0, B)
    total_ms, stats = pm.step_profiled(1, args.profile_steps)
    for name, s in stats.items():
        log(f"  {name}: {s['launches']} launches, mean {s['mean_ms'] * 1e3:.1f} us, "
            f"{s['bytes_per_launch'] / 1e6:.1f} MB alg")
    dom = max(stats.items(), key=lambda kv: kv[1]["mean_ms"] * kv[1]["launches"])
    achieved = dom[1]["bytes_per_launch"] / (dom[1]["mean_ms"] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom[0], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(dom[0], B),
            "alg_bytes_per_launch": dom[1]["bytes_per_launch"],
            "mean_launch_us": round(dom[1]["mean_ms"] * 1e3, 2)}
    pm.close()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_seconds)
        log(f"cpu baseline: {cpu['value']:.2f} Mcell*steps/s on {cpu['cores']} threads")

    out = {
        "metric": "Mcell*steps/s mixed-prec vs FP64 at <=1e-5 rel L2 (3D split-step)",
        "value": round(vm, 1),
        "unit": "Mcell*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tm * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 storage, fp64 arithmetic",
        "data": "synthetic: Gaussian beam + device-evaluated CSG sphere",
        "config": {"workload": "FFTSolver3D 512x512 over 20x20 um, 1000 z-steps of 10 nm, "
                               "lambda 0.1 nm, Gaussian waist 3 um, sphere r 4 um d 5e-6 b 1e-7",
                   "batch_fields": B, "grid": N, "z_steps": args.steps},
        "rel_l2_field_vs_fp64": ef,
        "rel_l2_intensity_vs_fp64": ei,
        "fp64_single": {"value": round(v64, 1), "unit": "Mcell*steps/s",
                        "ms_per_step": round(t64 * 1e3 / args.steps, 4)},
        "speedup_vs_fp64_single": round(vm / v64, 2),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(vm / cpu["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
