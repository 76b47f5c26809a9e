"""ctypes binding of libpaxgpu.so (include/paxgpu.h).

Fields cross the boundary as numpy complex128 arrays in the reference's
Armadillo column-major layout: a 3D field is an (Ny, Nx) Fortran-ordered
matrix (rows = y), a 2D field is a length-Nx vector.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpaxgpu.so")
_LIB = None

OK, ERR_INVALID, ERR_UNSUPPORTED, ERR_BUDGET, ERR_DEVICE, ERR_STATE, ERR_CANCELLED, ERR_INJECTED = range(8)
FFT2D, FFT3D, CN, ADI = 0, 1, 2, 3
PREC_FP64, PREC_MIXED = 0, 1
BC_DIRICHLET, BC_TRANSPARENT = 0, 1
FIX_ADI_Y_DIAG, FIX_REFR_SAMPLES, FIX_LINEAR_PAD = 1, 2, 4
LAYOUT_CONTIGUOUS, LAYOUT_INTERLEAVED = 0, 1
SHAPE_SPHERE, SHAPE_BOX, SHAPE_CYLINDER = 0, 1, 2
OP_UNION, OP_DIFFERENCE = 0, 1

SHAPE_DTYPE = np.dtype([("type", np.int32), ("op", np.int32), ("T", np.float64, 16),
                        ("p", np.float64, 3)])
PART_DTYPE = np.dtype([("first_shape", np.int32), ("n_shapes", np.int32), ("op", np.int32),
                       ("_pad", np.int32), ("delta", np.float64), ("beta", np.float64)])
MODULE_DTYPE = np.dtype([("first_part", np.int32), ("n_parts", np.int32)])

EXPORTS = [
    "pxg_ctx_create", "pxg_ctx_destroy", "pxg_ctx_reserved", "pxg_last_error", "pxg_version",
    "pxg_plan_create", "pxg_plan_destroy", "pxg_plan_bytes", "pxg_set_wavenumbers",
    "pxg_set_field", "pxg_get_field", "pxg_get_solution", "pxg_solution_dims",
    "pxg_set_material_csg", "pxg_set_material_curved_wg", "pxg_set_material_planes",
    "pxg_set_material_none", "pxg_step", "pxg_step_group", "pxg_sync", "pxg_step_profiled",
    "pxg_field_device_ptr", "pxg_thomas_batched", "pxg_plan_create_jobs",
    "pxg_set_job_material_curved_wg", "pxg_job_solution_dims", "pxg_plan_reset",
    "pxg_plan_set_lines", "pxg_plan_get_lines", "pxg_plan_autotune",
    "pxg_server_create", "pxg_server_submit", "pxg_server_status", "pxg_server_wait",
    "pxg_server_cancel", "pxg_server_stats_get", "pxg_server_destroy",
    "pxg_far_field_dims", "pxg_far_field", "pxg_intensity_trapezoid", "pxg_plan_set_graphs",
    "pxg_field_quantity_dims", "pxg_field_quantity", "pxg_field_uint8", "pxg_far_field_padded",
    "pxg_thomas_batched_noalloc", "pxg_set_equation_planes", "pxg_server_set_grouping",
    "pxg_plan_set_adi_tiles", "pxg_plan_get_adi_tiles",
]
THOMAS_NOALLOC_MAX_ROWS = 1 << 20
JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED, JOB_CANCELLED = range(5)
INJECT_NONE, INJECT_CREATE, INJECT_RESET = range(3)
MAT_NONE, MAT_CSG, MAT_CURVED_WG = 0, 1, 2
SRC_FIELD, SRC_SOLUTION = 0, 1
Q_ABS, Q_ABS2, Q_ARG, Q_REAL = 0, 1, 2, 3
PAD_ZERO, PAD_LINEAR_FIT, PAD_CONSTANT = 0, 1, 2


class PxgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"paxgpu error {code}: {msg}")
        self.code = code


class Problem(C.Structure):
    _fields_ = [("kind", C.c_int), ("precision", C.c_int), ("nx", C.c_int), ("nz", C.c_int),
                ("batch", C.c_int), ("xmin", C.c_double), ("dx", C.c_double),
                ("ymin", C.c_double), ("dy", C.c_double), ("zmin", C.c_double),
                ("dz", C.c_double), ("k", C.c_double), ("bc", C.c_int), ("use_tbc", C.c_int),
                ("absorb_npix_x", C.c_uint), ("absorb_npix_y", C.c_uint),
                ("absorb_inv_x", C.c_double), ("absorb_inv_y", C.c_double),
                ("quirks", C.c_uint), ("n_refr", C.c_uint), ("store_dsx", C.c_int),
                ("store_dsz", C.c_int)]


class CurvedWG(C.Structure):
    _fields_ = [("R", C.c_double), ("width", C.c_double), ("clad_delta", C.c_double),
                ("clad_beta", C.c_double), ("in_delta", C.c_double), ("in_beta", C.c_double),
                ("wglength", C.c_double), ("dx", C.c_double), ("dz", C.c_double),
                ("use_smoothed", C.c_int)]


class Job(C.Structure):
    _fields_ = [("prob", Problem), ("material", C.c_int), ("shapes", C.c_void_p),
                ("n_shapes", C.c_int), ("parts", C.c_void_p), ("n_parts", C.c_int),
                ("modules", C.c_void_p), ("n_modules", C.c_int), ("wg", CurvedWG),
                ("field", C.c_void_p), ("result", C.c_void_p), ("first_step", C.c_uint),
                ("nsteps", C.c_uint), ("inject_failure_after_chunks", C.c_int),
                ("inject_plan_failure", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("submitted", "completed", "failed", "cancelled",
                                          "rejected", "running", "queued")] + \
               [(n, C.c_double) for n in ("p50_ms", "p99_ms", "mean_ms", "max_ms", "cell_steps",
                                          "busy_s")] + \
               [("hbm_reserved", C.c_size_t), ("hbm_budget", C.c_size_t), ("chunks", C.c_uint64),
                ("grouped_chunks", C.c_uint64)]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char_p), ("mean_ms", C.c_double), ("launches", C.c_int),
                ("bytes_per_launch", C.c_double)]


def lib():
    """Load libpaxgpu.so; raises if it has not been built (no CPU fallback)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C paxpro_amd`")
        L = C.CDLL(LIB_PATH)
        P, I, U, D, S = C.c_void_p, C.c_int, C.c_uint, C.c_double, C.c_size_t
        L.pxg_last_error.restype = C.c_char_p
        L.pxg_version.restype = C.c_char_p
        L.pxg_ctx_create.argtypes = [I, S, P]
        L.pxg_ctx_destroy.argtypes = [P]
        L.pxg_ctx_reserved.restype = S
        L.pxg_ctx_reserved.argtypes = [P]
        L.pxg_plan_create.argtypes = [P, P, P]
        L.pxg_plan_destroy.argtypes = [P]
        L.pxg_plan_bytes.argtypes = [P, P]
        L.pxg_set_wavenumbers.argtypes = [P, P, I]
        L.pxg_set_field.argtypes = [P, I, P]
        L.pxg_get_field.argtypes = [P, I, P]
        L.pxg_get_solution.argtypes = [P, I, P, S]
        L.pxg_solution_dims.argtypes = [P, P]
        L.pxg_set_material_csg.argtypes = [P, P, I, P, I, P, I, I]
        L.pxg_set_material_curved_wg.argtypes = [P, P]
        L.pxg_set_material_planes.argtypes = [P, P, I, I, I]
        L.pxg_set_material_none.argtypes = [P]
        L.pxg_set_equation_planes.argtypes = [P, P, I, I, I]
        L.pxg_step.argtypes = [P, U, U]
        L.pxg_step_group.argtypes = [P, I, P, U]
        L.pxg_sync.argtypes = [P]
        L.pxg_step_profiled.argtypes = [P, U, U, P, I, P, P]
        L.pxg_field_device_ptr.argtypes = [P, I, P]
        L.pxg_thomas_batched.argtypes = [P, I, I, I, P, P, P]
        L.pxg_thomas_batched_noalloc.argtypes = [P, I, I, I, P, P, P, P]
        L.pxg_plan_create_jobs.argtypes = [P, P, I, P]
        L.pxg_plan_reset.argtypes = [P, P]
        L.pxg_plan_set_lines.argtypes = [P, I, I]
        L.pxg_plan_get_lines.argtypes = [P, P, P]
        L.pxg_plan_autotune.argtypes = [P, I]
        L.pxg_plan_set_adi_tiles.argtypes = [P, I, I]
        L.pxg_plan_get_adi_tiles.argtypes = [P, P, P]
        L.pxg_server_create.argtypes = [P, I, S, U, P]
        L.pxg_server_submit.argtypes = [P, P, P]
        L.pxg_server_status.argtypes = [P, C.c_uint64, P]
        L.pxg_server_wait.argtypes = [P, C.c_uint64, P, D]
        L.pxg_server_cancel.argtypes = [P, C.c_uint64]
        L.pxg_server_stats_get.argtypes = [P, P]
        L.pxg_server_destroy.argtypes = [P]
        L.pxg_server_set_grouping.argtypes = [P, I, I]
        L.pxg_set_job_material_curved_wg.argtypes = [P, I, P]
        L.pxg_job_solution_dims.argtypes = [P, I, P]
        L.pxg_far_field_dims.argtypes = [P, I, U, D, D, U, U, P]
        L.pxg_far_field.argtypes = [P, I, P, I, U, D, D, U, U, P]
        L.pxg_far_field_padded.argtypes = [P, I, P, I, U, I, D, D, U, U, P]
        L.pxg_intensity_trapezoid.argtypes = [P, I, I, P, P, P, P]
        L.pxg_field_quantity_dims.argtypes = [P, I, I, U, U, P]
        L.pxg_field_quantity.argtypes = [P, I, I, I, U, U, P]
        L.pxg_field_uint8.argtypes = [P, I, I, I, P, P, P]
        L.pxg_plan_set_graphs.argtypes = [P, I]
        _LIB = L
    return _LIB


def _check(rc):
    if rc != OK:
        raise PxgError(rc, lib().pxg_last_error().decode())


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def version():
    return lib().pxg_version().decode()


def problem(kind, nx, nz, dx, dz, k, xmin=0.0, ymin=None, dy=None, zmin=0.0, precision=PREC_FP64,
            batch=1, bc=BC_TRANSPARENT, use_tbc=1, absorb=(0, 0, 0.0, 0.0), quirks=0, n_refr=10,
            store_dsx=0, store_dsz=0):
    if ymin is None:
        ymin = xmin
    if dy is None:
        dy = dx
    return Problem(kind, precision, nx, nz, batch, xmin, dx, ymin, dy, zmin, dz, k, bc, use_tbc,
                   absorb[0], absorb[1], absorb[2], absorb[3], quirks, n_refr, store_dsx,
                   store_dsz)


class Ctx:
    def __init__(self, device=0, budget=0):
        self.h = C.c_void_p()
        _check(lib().pxg_ctx_create(device, budget, C.byref(self.h)))

    def reserved(self):
        return lib().pxg_ctx_reserved(self.h)

    def thomas_batched(self, diag, sub, rhs, layout=LAYOUT_CONTIGUOUS):
        """ThomasAlgorithm::solve for nsys systems: arrays (nsys, n) / (nsys, n-1)
        complex128 for the contiguous layout, (n, nsys) / (n-1, nsys) interleaved.
        Returns (solution, clobbered rhs)."""
        d = np.ascontiguousarray(diag, dtype=np.complex128).copy()
        s = np.ascontiguousarray(sub, dtype=np.complex128)
        r = np.ascontiguousarray(rhs, dtype=np.complex128).copy()
        if layout == LAYOUT_CONTIGUOUS:
            nsys, n = d.shape
        else:
            n, nsys = d.shape
        _check(lib().pxg_thomas_batched(self.h, layout, nsys, n, _p(d), _p(s), _p(r)))
        return d, r

    def thomas_batched_noalloc(self, d_diag, d_sub, d_rhs, nsys, n, layout=LAYOUT_CONTIGUOUS,
                               stream=None):
        """pxg_thomas_batched_noalloc on device addresses (ints): enqueues the
        solve on `stream` (None = the null stream) and returns without waiting."""
        _check(lib().pxg_thomas_batched_noalloc(self.h, layout, nsys, n, d_diag, d_sub, d_rhs, stream))

    def close(self):
        if self.h:
            _check(lib().pxg_ctx_destroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plan_bytes(prob):
    n = C.c_size_t()
    _check(lib().pxg_plan_bytes(C.byref(prob), C.byref(n)))
    return n.value


MAX_GROUP = 8
ADI_TILES = 6
SERVER_GROUP_MAX_NX = 512   # the server's default: larger grids fill the device alone


def step_group(plans, first_steps, nsteps):
    """pxg_step_group: advance batch-1 FFT3D plans of one nx and precision in
    shared launches, plan i from first_steps[i] (an int applies to all)."""
    if isinstance(first_steps, int):
        first_steps = [first_steps] * len(plans)
    hs = (C.c_void_p * len(plans))(*[p.h.value for p in plans])
    fs = (C.c_uint * len(plans))(*first_steps)
    _check(lib().pxg_step_group(C.cast(hs, C.c_void_p), len(plans), C.cast(fs, C.c_void_p), nsteps))


class Plan:
    def __init__(self, ctx, prob):
        self.ctx = ctx
        self.prob = prob
        self.h = C.c_void_p()
        _check(lib().pxg_plan_create(ctx.h, C.byref(prob), C.byref(self.h)))
        self.is3d = prob.kind in (FFT3D, ADI)
        self.nx = prob.nx

    def shape(self):
        return (self.nx, self.nx) if self.is3d else (self.nx,)

    def reset(self, prob):
        _check(lib().pxg_plan_reset(self.h, C.byref(prob)))
        self.prob = prob

    def set_lines(self, lines_x=0, lines_y=0):
        """Lines per work-group of the FFT3D x and y passes (0 keeps one)."""
        _check(lib().pxg_plan_set_lines(self.h, lines_x, lines_y))

    def lines(self):
        lx, ly = C.c_int(), C.c_int()
        _check(lib().pxg_plan_get_lines(self.h, C.byref(lx), C.byref(ly)))
        return lx.value, ly.value

    def set_adi_tiles(self, x_tile=-1, y_tile=-1):
        """Tile shapes of the ADI x- and y-implicit half-steps (-1 keeps one)."""
        _check(lib().pxg_plan_set_adi_tiles(self.h, x_tile, y_tile))

    def adi_tiles(self):
        tx, ty = C.c_int(), C.c_int()
        _check(lib().pxg_plan_get_adi_tiles(self.h, C.byref(tx), C.byref(ty)))
        return tx.value, ty.value

    def autotune(self, reps=0):
        """Time every supported work-group shape (FFT3D) or tile shape (ADI) and
        keep the fastest; zeroes the field, so call it before set_field."""
        _check(lib().pxg_plan_autotune(self.h, reps))
        return self.adi_tiles() if self.prob.kind == ADI else self.lines()

    def set_field(self, field, b=0):
        order = "F" if self.is3d else "C"
        f = np.asarray(field, dtype=np.complex128)
        if f.shape != self.shape():
            raise ValueError(f"field shape {f.shape} != {self.shape()}")
        f = np.require(f, requirements=[order[0] if order == "F" else "C"]).copy(order=order)
        _check(lib().pxg_set_field(self.h, b, _p(f)))

    def get_field(self, b=0):
        out = np.empty(self.shape(), dtype=np.complex128, order="F" if self.is3d else "C")
        _check(lib().pxg_get_field(self.h, b, _p(out)))
        return out

    def set_wavenumbers(self, ks):
        ks = np.ascontiguousarray(ks, dtype=np.float64)
        _check(lib().pxg_set_wavenumbers(self.h, _p(ks), ks.size))

    def set_graphs(self, enable=True):
        """Capture FFT3D z-step loops as HIP graphs and replay them (pxg_plan_set_graphs)."""
        _check(lib().pxg_plan_set_graphs(self.h, int(enable)))

    def intensity_trapezoid(self, steps, ix_start, ix_end, b=0):
        """Trapezoid sums of |u|^2 over [ix_start[i], ix_end[i]] of stored solution column
        steps[i] (post::Transmittivity::trapezoidalIntegrateIntensityZ without its 0.5 dx)."""
        st = np.ascontiguousarray(steps, dtype=np.int32)
        a = np.ascontiguousarray(ix_start, dtype=np.int32)
        e = np.ascontiguousarray(ix_end, dtype=np.int32)
        out = np.empty(st.size, dtype=np.float64)
        _check(lib().pxg_intensity_trapezoid(self.h, b, st.size, _p(st), _p(a), _p(e), _p(out)))
        return out

    def field_quantity(self, quantity, source=SRC_FIELD, b=0, export_rows=0, export_cols=0):
        """|u|, |u|^2, arg u or Re u of the current field or the stored solution, with
        resizeMatrix for matrices (post::Intensity, ExitIntensity, Phase, ExitPhase,
        ExitField; src/postProcessMod.cpp:7-151).  Column-major (rows, cols, slices)."""
        dims = np.zeros(3, dtype=np.int32)
        _check(lib().pxg_field_quantity_dims(self.h, b, source, export_rows, export_cols, _p(dims)))
        out = np.empty(tuple(int(d) for d in dims), dtype=np.float64, order="F")
        _check(lib().pxg_field_quantity(self.h, b, source, quantity, export_rows, export_cols, _p(out)))
        return out

    def field_uint8(self, source=SRC_SOLUTION, b=0, log_scale=False, min_value=None, max_value=None):
        """8-bit |u| image (post::IntensityUint8) or log|u| image (post::LogIntensityUint8,
        min_value / max_value as given to setMinValue / setMaxValue)."""
        dims = np.zeros(3, dtype=np.int32)
        _check(lib().pxg_field_quantity_dims(self.h, b, source, 0, 0, _p(dims)))
        out = np.empty(tuple(int(d) for d in dims), dtype=np.uint8, order="F")
        lo = None if min_value is None else np.array([min_value], dtype=np.float64)
        hi = None if max_value is None else np.array([max_value], dtype=np.float64)
        _check(lib().pxg_field_uint8(self.h, b, source, int(bool(log_scale)),
                                     None if lo is None else _p(lo), None if hi is None else _p(hi),
                                     _p(out)))
        return out

    def far_field(self, b=0, ref=None, ref_b=0, pad=1024, phi_min=-90.0, phi_max=90.0,
                  export_rows=0, export_cols=0, padding=PAD_ZERO):
        """Scattered far-field intensity of field b minus field ref_b of plan `ref`
        (post::FarField::result, src/farFieldPost.cpp:141-230): rows follow y, columns x;
        padding PAD_ZERO or PAD_CONSTANT (post::FarField::setPadding)."""
        dims = np.zeros(2, dtype=np.int32)
        _check(lib().pxg_far_field_dims(self.h, b, pad, phi_min, phi_max, export_rows,
                                        export_cols, _p(dims)))
        out = np.empty((int(dims[0]), int(dims[1])), dtype=np.float64, order="F")
        _check(lib().pxg_far_field_padded(self.h, b, None if ref is None else ref.h, ref_b, pad,
                                          padding, phi_min, phi_max, export_rows, export_cols,
                                          _p(out)))
        return out

    def equation_planes(self, coef, first_plane=0, z_invariant=False):
        """ParaxialEquation planes (n_planes, nx, 5): F, G, H(x+dx/2), H(x-dx/2), J; None
        restores the default equation."""
        if coef is None:
            _check(lib().pxg_set_equation_planes(self.h, None, 0, 0, 0))
            return
        c = np.ascontiguousarray(coef, dtype=np.float64)
        _check(lib().pxg_set_equation_planes(self.h, _p(c), c.shape[0], first_plane, int(z_invariant)))

    def material_none(self):
        _check(lib().pxg_set_material_none(self.h))

    def material_csg(self, shapes, parts, modules, reference_run=False):
        self._csg = [np.ascontiguousarray(shapes, SHAPE_DTYPE), np.ascontiguousarray(parts, PART_DTYPE),
                     np.ascontiguousarray(modules, MODULE_DTYPE)]
        s, p, m = self._csg
        _check(lib().pxg_set_material_csg(self.h, _p(s), len(s), _p(p), len(p), _p(m), len(m),
                                          int(reference_run)))

    def material_curved(self, **kw):
        d = dict(in_delta=0.0, in_beta=0.0, wglength=1e10, use_smoothed=0)
        d.update(kw)
        wg = CurvedWG(**d)
        _check(lib().pxg_set_material_curved_wg(self.h, C.byref(wg)))

    def material_planes(self, db, first_plane=0, z_invariant=False):
        """db: (n_planes, cells, 2) float64, cells in the field's memory order."""
        db = np.ascontiguousarray(db, dtype=np.float64)
        _check(lib().pxg_set_material_planes(self.h, _p(db), db.shape[0], first_plane,
                                             int(z_invariant)))

    def step(self, first, nsteps):
        _check(lib().pxg_step(self.h, first, nsteps))

    def sync(self):
        _check(lib().pxg_sync(self.h))

    def step_profiled(self, first, nsteps, max_stats=16):
        stats = (KernelStat * max_stats)()
        n = C.c_int()
        total = C.c_double()
        _check(lib().pxg_step_profiled(self.h, first, nsteps, stats, max_stats, C.byref(n),
                                       C.byref(total)))
        out = {}
        for i in range(n.value):
            s = stats[i]
            out[s.name.decode()] = dict(mean_ms=s.mean_ms, launches=s.launches,
                                        bytes_per_launch=s.bytes_per_launch)
        return total.value, out

    def solution_dims(self):
        d = (C.c_int * 3)()
        _check(lib().pxg_solution_dims(self.h, d))
        return tuple(d)

    def get_solution(self, b=0):
        d = self.solution_dims()
        if self.is3d:
            out = np.empty(d, dtype=np.complex128, order="F")
        else:
            out = np.empty((d[0], d[1]), dtype=np.complex128, order="F")
        _check(lib().pxg_get_solution(self.h, b, _p(out), out.size))
        return out

    def device_ptr(self, b=0):
        p = C.c_void_p()
        _check(lib().pxg_field_device_ptr(self.h, b, C.byref(p)))
        return p.value

    def close(self):
        if self.h:
            _check(lib().pxg_plan_destroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class JobPlan(Plan):
    """Independent CN problems with their own grids in one plan (pxg_plan_create_jobs)."""

    def __init__(self, ctx, probs):
        self.ctx = ctx
        self.probs = list(probs)
        arr = (Problem * len(self.probs))(*self.probs)
        self.h = C.c_void_p()
        _check(lib().pxg_plan_create_jobs(ctx.h, arr, len(self.probs), C.byref(self.h)))
        self.is3d = False
        self.prob = self.probs[0]

    def shape_of(self, b):
        return (self.probs[b].nx,)

    def set_field(self, field, b=0):
        f = np.ascontiguousarray(field, dtype=np.complex128)
        if f.shape != self.shape_of(b):
            raise ValueError(f"field shape {f.shape} != {self.shape_of(b)}")
        _check(lib().pxg_set_field(self.h, b, _p(f)))

    def get_field(self, b=0):
        out = np.empty(self.shape_of(b), dtype=np.complex128)
        _check(lib().pxg_get_field(self.h, b, _p(out)))
        return out

    def material_curved_job(self, b, **kw):
        d = dict(in_delta=0.0, in_beta=0.0, wglength=1e10, use_smoothed=0)
        d.update(kw)
        wg = CurvedWG(**d)
        _check(lib().pxg_set_job_material_curved_wg(self.h, b, C.byref(wg)))

    def get_solution(self, b=0):
        d = (C.c_int * 3)()
        _check(lib().pxg_job_solution_dims(self.h, b, d))
        out = np.empty((d[0], d[1]), dtype=np.complex128, order="F")
        _check(lib().pxg_get_solution(self.h, b, _p(out), out.size))
        return out


class Server:
    """pxg_server: persistent request loop with HBM-budget admission."""

    def __init__(self, ctx, slots=4, budget=0, chunk_steps=32, max_group=None, group_max_nx=None):
        self.ctx = ctx
        self.h = C.c_void_p()
        _check(lib().pxg_server_create(ctx.h, slots, budget, chunk_steps, C.byref(self.h)))
        self._keep = {}
        if max_group is not None or group_max_nx is not None:
            self.set_grouping(MAX_GROUP if max_group is None else max_group,
                              SERVER_GROUP_MAX_NX if group_max_nx is None else group_max_nx)

    def set_grouping(self, max_group, max_nx=None):
        _check(lib().pxg_server_set_grouping(self.h, max_group, SERVER_GROUP_MAX_NX if max_nx is None else max_nx))

    def submit(self, prob, field, nsteps, first_step=1, csg=None, curved=None, inject=-1,
               inject_plan=INJECT_NONE):
        is3d = prob.kind in (FFT3D, ADI)
        f = np.asarray(field, dtype=np.complex128)
        f = np.asfortranarray(f) if is3d else np.ascontiguousarray(f)
        out = np.empty_like(f, order="F" if is3d else "C")
        job = Job()
        job.prob = prob
        keep = [f, out]
        if csg is not None:
            s, pa, m = (np.ascontiguousarray(csg[0], SHAPE_DTYPE), np.ascontiguousarray(csg[1], PART_DTYPE),
                        np.ascontiguousarray(csg[2], MODULE_DTYPE))
            keep += [s, pa, m]
            job.material = MAT_CSG
            job.shapes, job.n_shapes = _p(s), len(s)
            job.parts, job.n_parts = _p(pa), len(pa)
            job.modules, job.n_modules = _p(m), len(m)
        elif curved is not None:
            d = dict(in_delta=0.0, in_beta=0.0, wglength=1e10, use_smoothed=0)
            d.update(curved)
            job.material = MAT_CURVED_WG
            job.wg = CurvedWG(**d)
        else:
            job.material = MAT_NONE
        job.field, job.result = _p(f), _p(out)
        job.first_step, job.nsteps = first_step, nsteps
        job.inject_failure_after_chunks = inject
        job.inject_plan_failure = inject_plan
        jid = C.c_uint64()
        _check(lib().pxg_server_submit(self.h, C.byref(job), C.byref(jid)))
        self._keep[jid.value] = (keep, out)
        return jid.value

    def status(self, jid):
        s = C.c_int()
        _check(lib().pxg_server_status(self.h, jid, C.byref(s)))
        return s.value

    def wait(self, jid, timeout=0.0):
        """Returns (status code, latency ms, result array)."""
        lat = C.c_double()
        rc = lib().pxg_server_wait(self.h, jid, C.byref(lat), timeout)
        keep, out = self._keep.pop(jid)
        return rc, lat.value, out

    def cancel(self, jid):
        _check(lib().pxg_server_cancel(self.h, jid))

    def stats(self):
        st = ServerStats()
        _check(lib().pxg_server_stats_get(self.h, C.byref(st)))
        return {n: getattr(st, n) for n, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            _check(lib().pxg_server_destroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
