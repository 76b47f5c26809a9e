"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check of
__graft_entry__.py and by the cpu_baseline leg of bench.py, always as the
checker (or the timed CPU baseline), never as the product path.

Arrays are numpy complex128 in the reference's Armadillo column-major layout
(3D field: Ny rows x Nx cols, element (r, c) at c*Ny + r; numpy arrays are
passed as shape (Nx, Ny) C-order == (Ny, Nx) Fortran-order).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

SHAPE_DTYPE = np.dtype([("type", np.int32), ("op", np.int32), ("T", np.float64, 16),
                        ("p", np.float64, 3)])
PART_DTYPE = np.dtype([("first_shape", np.int32), ("n_shapes", np.int32), ("op", np.int32),
                       ("_pad", np.int32), ("delta", np.float64), ("beta", np.float64)])
MODULE_DTYPE = np.dtype([("first_part", np.int32), ("n_parts", np.int32)])

SPHERE, BOX, CYLINDER = 0, 1, 2
UNION, DIFFERENCE = 0, 1
FIX_ADI_Y_DIAG = 1
FIX_REFR_SAMPLES = 2
FIX_LINEAR_PAD = 4
BC_DIRICHLET, BC_TRANSPARENT = 0, 1


class Grid(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("xmin", "xmax", "dx", "ymin", "ymax", "dy", "zmin", "zmax", "dz", "k")]


class Material(C.Structure):
    _fields_ = [("fn", C.c_void_p), ("ctx", C.c_void_p)]


class CSG(C.Structure):
    _fields_ = [("shapes", C.c_void_p), ("parts", C.c_void_p), ("modules", C.c_void_p),
                ("n_modules", C.c_int), ("is_reference_run", C.c_int)]


class CurvedWG(C.Structure):
    _fields_ = [("R", C.c_double), ("width", C.c_double), ("clad_delta", C.c_double),
                ("clad_beta", C.c_double), ("in_delta", C.c_double), ("in_beta", C.c_double),
                ("wglength", C.c_double), ("dx", C.c_double), ("dz", C.c_double),
                ("use_smoothed", C.c_int)]


class Slab(C.Structure):
    _fields_ = [("x_lo", C.c_double), ("x_hi", C.c_double), ("core_delta", C.c_double),
                ("core_beta", C.c_double), ("clad_delta", C.c_double), ("clad_beta", C.c_double)]


class Equation(C.Structure):
    _fields_ = [("fn", C.c_void_p), ("ctx", C.c_void_p)]


class PolyEq:
    """A ParaxialEquation whose F, G, H, J are each c0 + cx x + cz z + cinv / x
    (rows of `coef`, shape (4, 4)); evaluated by the oracle's orc_poly_eqfn."""

    def __init__(self, coef):
        self.coef = np.ascontiguousarray(coef, dtype=np.float64).reshape(4, 4)
        self.eq = Equation(C.cast(lib().orc_poly_eqfn, C.c_void_p), self.coef.ctypes.data_as(C.c_void_p))

    def values(self, x, z):
        """(F, G, H, J) at (x, z), as the oracle evaluates them."""
        out = np.zeros(4)
        lib().orc_poly_eqfn(self.coef.ctypes.data_as(C.c_void_p), x, z, _p(out))
        return out


class Absorber3D(C.Structure):
    _fields_ = [("npix_x", C.c_uint), ("npix_y", C.c_uint), ("inv_x", C.c_double),
                ("inv_y", C.c_double)]


class Gauss(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("k", "amplitude", "z0", "waist", "cx", "cy", "angle_x_deg", "angle_y_deg")]


class Plane(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("k", "amplitude", "z0", "angle_rad")]


def lib():
    global _LIB
    if _LIB is None:
        # PAXORACLE_LIB names another build of paxoracle.c (the mutation
        # tests load deliberately broken copies through it)
        path = os.environ.get("PAXORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(path)
        P = C.c_void_p
        L.orc_nodes.restype = C.c_uint
        L.orc_nodes.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_fft1d.argtypes = [P, C.c_uint, C.c_int]
        L.orc_fft2d.argtypes = [P, C.c_uint, C.c_uint, C.c_int]
        L.orc_thomas.argtypes = [P, P, P, C.c_uint]
        for f in ("orc_shape_translate",):
            getattr(L, f).argtypes = [P, C.c_double, C.c_double, C.c_double]
        L.orc_shape_rotate.argtypes = [P, C.c_double, C.c_int]
        L.orc_shape_scale.argtypes = [P, C.c_double, C.c_int]
        L.orc_mat4_identity.argtypes = [P]
        L.orc_shape_transform.argtypes = [P, P, P, P]
        L.orc_csg_eval.argtypes = [P, C.c_double, C.c_double, C.c_double, P, P]
        L.orc_fft3d_run.argtypes = [P, P, P, C.c_uint, C.c_uint, P, C.c_uint, C.c_uint, C.c_uint,
                                    P, C.c_uint, C.c_uint, C.c_uint]
        L.orc_cube_store.argtypes = [P, C.c_uint, P, C.c_uint, C.c_uint, C.c_uint, C.c_uint]
        L.orc_fft2d_run.argtypes = [P, P, P, C.c_uint, C.c_uint, C.c_uint, P, C.c_uint]
        L.orc_cn_run.argtypes = [P, P, C.c_int, P, C.c_uint, C.c_uint, C.c_uint, P, C.c_uint,
                                 C.c_uint, C.c_double]
        L.orc_cn_step_planes.argtypes = [C.c_double, C.c_double, C.c_int, P, P, C.c_uint,
                                         P, P, P, P]
        L.orc_cn_run_eq.argtypes = [P, P, P, C.c_int, P, C.c_uint, C.c_uint, C.c_uint, P, C.c_uint,
                                    C.c_uint, C.c_double]
        L.orc_poly_eqfn.argtypes = [P, C.c_double, C.c_double, P]
        L.orc_adi_run.argtypes = [P, P, C.c_int, C.c_uint, P, C.c_uint, C.c_uint, C.c_uint,
                                  P, C.c_uint, C.c_uint, C.c_uint]
        L.orc_gauss_2d_v.argtypes = [P, C.c_double, C.c_double, P]
        L.orc_plane_wave_v.argtypes = [P, C.c_double, C.c_double, P]
        L.orc_gauss_fill3d.argtypes = [P, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_uint, P]
        L.orc_num_threads.restype = C.c_int
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def nodes(vmin, vmax, step):
    return lib().orc_nodes(vmin, vmax, step)


def fft1d(a, sign=-1):
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    lib().orc_fft1d(_p(a), a.size, sign)
    return a


def fft2d_colmajor(a, sign=-1):
    """2D DFT of a Fortran-ordered (nrows, ncols) complex matrix."""
    a = np.asfortranarray(a, dtype=np.complex128).copy(order="F")
    lib().orc_fft2d(_p(a), a.shape[0], a.shape[1], sign)
    return a


def thomas(diag, sub, rhs):
    d = np.ascontiguousarray(diag, dtype=np.complex128).copy()
    s = np.ascontiguousarray(sub, dtype=np.complex128).copy()
    r = np.ascontiguousarray(rhs, dtype=np.complex128).copy()
    lib().orc_thomas(_p(d), _p(s), _p(r), d.size)
    return d


# --------------------------------------------------------------------------- CSG
def mat4_identity():
    T = np.zeros(16)
    lib().orc_mat4_identity(_p(T))
    return T


def translate(T, x, y, z):
    lib().orc_shape_translate(_p(T), x, y, z)


def rotate(T, deg, axis):
    lib().orc_shape_rotate(_p(T), deg, axis)


def scale(T, f, axis):
    lib().orc_shape_scale(_p(T), f, axis)


def transform(T, x, y, z):
    xs, ys, zs = (C.c_double(v) for v in (x, y, z))
    lib().orc_shape_transform(_p(T), C.byref(xs), C.byref(ys), C.byref(zs))
    return xs.value, ys.value, zs.value


class CSGMaterial:
    """Holds the flat CSG arrays alive and exposes an orc_material."""

    def __init__(self, shapes, parts, modules, reference_run=False):
        self.shapes = np.ascontiguousarray(shapes, dtype=SHAPE_DTYPE)
        self.parts = np.ascontiguousarray(parts, dtype=PART_DTYPE)
        self.modules = np.ascontiguousarray(modules, dtype=MODULE_DTYPE)
        self.csg = CSG(_p(self.shapes), _p(self.parts), _p(self.modules), len(self.modules),
                       int(reference_run))
        fn = C.cast(lib().orc_csg_matfn, C.c_void_p)
        self.mat = Material(fn, C.cast(C.byref(self.csg), C.c_void_p))

    def eval(self, x, y, z):
        d, b = C.c_double(), C.c_double()
        lib().orc_csg_eval(C.byref(self.csg), x, y, z, C.byref(d), C.byref(b))
        return d.value, b.value


class CurvedMaterial:
    def __init__(self, **kw):
        defaults = dict(in_delta=0.0, in_beta=0.0, wglength=1e10, use_smoothed=0)
        defaults.update(kw)
        self.wg = CurvedWG(**defaults)
        self.mat = Material(C.cast(lib().orc_curved_matfn, C.c_void_p),
                            C.cast(C.byref(self.wg), C.c_void_p))


class SlabMaterial:
    def __init__(self, x_lo, x_hi, core_delta, core_beta, clad_delta, clad_beta):
        self.slab = Slab(x_lo, x_hi, core_delta, core_beta, clad_delta, clad_beta)
        self.mat = Material(C.cast(lib().orc_slab_matfn, C.c_void_p),
                            C.cast(C.byref(self.slab), C.c_void_p))


class NoMaterial:
    mat = Material(None, None)


def grid(xmin, xmax, dx, zmin, zmax, dz, k, ymin=None, ymax=None, dy=None):
    if ymin is None:
        ymin, ymax, dy = xmin, xmax, dx
    return Grid(xmin, xmax, dx, ymin, ymax, dy, zmin, zmax, dz, k)


# ---------------------------------------------------------------------- sources
def gauss3d_field(g, gb, n):
    """ParaxialSimulation::setBoundaryConditions 3D: values(j,i) = src(x_i, y_j, 0)."""
    out = np.empty((n, n), dtype=np.complex128, order="F")  # (rows=y, cols=x)
    lib().orc_gauss_fill3d(C.byref(gb), g.xmin, g.dx, g.ymin, g.dy, n, _p(out))
    return out


def gauss2d_value(gb, x, z=0.0):
    v = np.zeros(2)
    lib().orc_gauss_2d_v(C.byref(gb), x, z, _p(v))
    return complex(v[0], v[1])


def plane_value(pw, x, z=0.0):
    v = np.zeros(2)
    lib().orc_plane_wave_v(C.byref(pw), x, z, _p(v))
    return complex(v[0], v[1])


# ---------------------------------------------------------------------- solvers
def fft3d_run(g, material, field, first_step, nsteps, absorber=None, n_refr=10, quirks=0,
              cube=None, dsz=1):
    f = np.asfortranarray(field, dtype=np.complex128).copy(order="F")
    n = f.shape[0]
    ab = absorber if absorber is not None else Absorber3D(0, 0, 0.0, 0.0)
    cube_p, cn, cs = (None, 0, 0) if cube is None else (_p(cube), cube.shape[0], cube.shape[2])
    lib().orc_fft3d_run(C.byref(g), C.byref(material.mat), C.byref(ab), n_refr, quirks, _p(f), n,
                        first_step, nsteps, cube_p, cn, cs, dsz)
    return f


def fft2d_run(g, material, field, first_step, nsteps, solution=None):
    f = np.ascontiguousarray(field, dtype=np.complex128).copy()
    sp, nz = (None, 0) if solution is None else (_p(solution), solution.shape[1])
    lib().orc_fft2d_run(C.byref(g), C.byref(material.mat), _p(f), f.size, first_step, nsteps,
                        sp, nz)
    return f


def cn_run(g, material, field, first_step, nsteps, bc=BC_TRANSPARENT, solution=None,
           absorb_npix=0, absorb_inv=0.0, equation=None):
    """equation: a PolyEq (any ParaxialEquation restated as orc_equation), or None
    for the default F = G = H = 1, J = 0."""
    f = np.ascontiguousarray(field, dtype=np.complex128).copy()
    sp, nz = (None, 0) if solution is None else (_p(solution), solution.shape[1])
    eq = None if equation is None else C.byref(equation.eq)
    lib().orc_cn_run_eq(C.byref(g), C.byref(material.mat), eq, bc, _p(f), f.size, first_step, nsteps,
                        sp, nz, absorb_npix, absorb_inv)
    return f


def cn_step_planes(rho, r, bc, prev, d1, b1, d0, b0):
    prev = np.ascontiguousarray(prev, dtype=np.complex128)
    cur = np.empty_like(prev)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (d1, b1, d0, b0)]
    lib().orc_cn_step_planes(rho, r, bc, _p(prev), _p(cur), prev.size, *[_p(a) for a in arrs])
    return cur


def adi_run(g, material, field, first_step, nsteps, use_tbc=True, quirks=0, cube=None, dsz=1):
    f = np.asfortranarray(field, dtype=np.complex128).copy(order="F")
    n = f.shape[0]
    cube_p, cn, cs = (None, 0, 0) if cube is None else (_p(cube), cube.shape[0], cube.shape[2])
    lib().orc_adi_run(C.byref(g), C.byref(material.mat), int(use_tbc), quirks, _p(f), n,
                      first_step, nsteps, cube_p, cn, cs, dsz)
    return f


def cube_store(field, cube, dsz, step):
    f = np.asfortranarray(field, dtype=np.complex128)
    lib().orc_cube_store(_p(f), f.shape[0], _p(cube), cube.shape[0], cube.shape[2], dsz, step)


def num_threads():
    return lib().orc_num_threads()


def rel_l2(a, ref):
    """Relative L2 distance.  Arrays of one shape are compared element by
    element whatever their memory order; otherwise both are flattened in their
    memory order (a flat column-major buffer against a Fortran matrix)."""
    a, ref = np.asarray(a), np.asarray(ref)
    if a.shape != ref.shape:
        a, ref = a.ravel(order="K"), ref.ravel(order="K")
    return float(np.linalg.norm(a - ref) / np.linalg.norm(ref))


def rel_l2_intensity(a, ref):
    ia = np.abs(np.asarray(a)) ** 2
    ir = np.abs(np.asarray(ref)) ** 2
    return float(np.linalg.norm(ia - ir) / np.linalg.norm(ir))


# ---------------------------------------------------------------------- far field
def far_field_angle_index(angle_deg, size, k, step):
    """post::FarField::farFieldAngleToIndx (src/farFieldPost.cpp:80-103): index of a
    scattering angle in the fftshifted spectrum of `size` points, spacing `step`."""
    pi = np.arccos(-1.0)
    v = size * k * step * np.sin(angle_deg * pi / 180.0) / (2.0 * pi)
    if abs(v) >= size // 2:  # covers values beyond the int range as well
        return size - 1 if angle_deg > 0.0 else 0
    n = int(v)  # C++ conversion: truncation toward zero
    if abs(n) >= size // 2:
        return size - 1 if angle_deg > 0.0 else 0
    return n + size // 2


def far_field_window(n, pad, k, dx, dy, phi_min, phi_max):
    """Signal length and the (row, column) index windows of the far field
    (src/farFieldPost.cpp:141-156,193-196): rows follow y (dy), columns x (dx)."""
    M = max(pad, n)
    ylo, yhi = far_field_angle_index(phi_min, M, k, dy), far_field_angle_index(phi_max, M, k, dy)
    xlo, xhi = far_field_angle_index(phi_min, M, k, dx), far_field_angle_index(phi_max, M, k, dx)
    return M, (ylo, yhi), (xlo, xhi)


def resize_matrix(orig, rows, cols):
    """PostProcessingModule::resizeMatrix (src/postProcessing.cpp:29-49): block means
    over inclusive, truncated index ranges (neighbouring blocks share a row or
    column); unchanged when either export size reaches the matrix size."""
    if rows >= orig.shape[0] or cols >= orig.shape[1]:
        return orig.copy()
    dX = orig.shape[1] / (cols + 1.0)
    dY = orig.shape[0] / (rows + 1.0)
    out = np.empty((rows, cols), order="F")
    for i in range(cols):
        for j in range(rows):
            blk = orig[int(j * dY):int((j + 1) * dY) + 1, int(i * dX):int((i + 1) * dX) + 1]
            out[j, i] = blk.mean(axis=0).mean()  # mean of the column means
    return out


def field_quantity(u, quantity, export_rows=0, export_cols=0):
    """The field quantities of the post-processing modules (src/postProcessMod.cpp:7-151):
    quantity 0 |u| (Intensity, :7-15), 1 |u|^2 (ExitIntensity, :122-136), 2 atan2(Im, Re)
    (Phase and ExitPhase, :90-104, :138-151), 3 Re u (ExitField, :106-120).  A matrix
    (2D array) is then resized by resizeMatrix when both export sizes are set; vectors
    and cubes are returned as they are."""
    u = np.asarray(u)
    if quantity == 0:
        r = np.abs(u)
    elif quantity == 1:
        r = np.abs(u) ** 2
    elif quantity == 2:
        r = np.arctan2(u.imag, u.real)
    elif quantity == 3:
        r = u.real.copy()
    else:
        raise ValueError(quantity)
    if r.ndim == 2 and export_rows and export_cols:
        r = resize_matrix(np.asfortranarray(r), export_rows, export_cols)
    return r


def _to_uint8(v):
    """Assignment of a double in [0, 256) to unsigned char: truncation.  Values the
    reference would convert with undefined behaviour (NaN) become 0."""
    v = np.asarray(v, dtype=np.float64)
    out = np.zeros(v.shape, dtype=np.uint8)
    ok = (v >= 0.0) & (v < 256.0)
    out[ok] = np.trunc(v[ok]).astype(np.uint8)
    return out


def intensity_uint8(u):
    """post::IntensityUint8::result (src/postProcessMod.cpp:17-30): 255 |u| / |max u|,
    where Armadillo's max() of a complex cube is the element of largest modulus."""
    a = np.abs(np.asarray(u))
    with np.errstate(divide="ignore", invalid="ignore"):
        return _to_uint8(255.0 * a / a.max())


def log_intensity_uint8(u, min_value=None, max_value=None):
    """post::LogIntensityUint8::result (src/postProcessMod.cpp:32-76): log|u| mapped from
    [bot, top] to [0, 255], clamped outside.  top = log|max u|, lowered to
    log(max_value) when setMaxValue was called and log|max u| exceeds it; bot =
    log|min u|, raised to log(min_value) when setMinValue was called and log|min u| is
    below it (setMinValue / setMaxValue store logarithms, :78-88)."""
    a = np.abs(np.asarray(u))
    with np.errstate(divide="ignore", invalid="ignore"):
        lmax, lmin = np.log(a.max()), np.log(a.min())
        top, bot = lmax, lmin
        if max_value is not None:
            hi = np.log(max_value)
            top = hi if lmax > hi else lmax
        if min_value is not None:
            lo = np.log(min_value)
            bot = lo if lmin < lo else lmin
        l = np.log(a)
        v = np.where(l > top, 255.0, np.where(l < bot, 0.0, 255.0 * (l - bot) / (top - bot)))
    return _to_uint8(v)


def _ref_fftshift(a, axis):
    """post::FarField::fftshift (src/farFieldPost.cpp:60-71): swaps element i with
    i + N/2 for i < N/2 (for odd N the last element stays)."""
    N = a.shape[axis]
    h = N // 2
    idx = np.arange(N)
    idx[:h], idx[h:2 * h] = np.arange(h, 2 * h), np.arange(h)
    return np.take(a, idx, axis=axis)


def _constant_pad(lines, M, axis):
    """post::FarField::constantPadding (src/farFieldPost.cpp:314-321): entries n .. M-1
    of every line (along `axis`) repeat the line's entry 0."""
    n = lines.shape[axis]
    shape = list(lines.shape)
    shape[axis] = M
    out = np.empty(shape, dtype=np.complex128)
    if axis == 0:
        out[:n] = lines
        out[n:] = lines[0:1, :]
    else:
        out[:, :n] = lines
        out[:, n:] = lines[:, 0:1]
    return out


def _gsl_fit_linear(x, y):
    """gsl_fit_linear (GSL fit/linear.c, the library the reference links; its published
    algorithm): running means of x and y, then running means of dx^2 and dx dy;
    c1 = m_dxdy / m_dx2, c0 = m_y - m_x c1."""
    mx = my = 0.0
    for i in range(len(x)):
        mx += (x[i] - mx) / (i + 1.0)
        my += (y[i] - my) / (i + 1.0)
    mdx2 = mdxdy = 0.0
    for i in range(len(x)):
        ddx, ddy = x[i] - mx, y[i] - my
        mdx2 += (ddx * ddx - mdx2) / (i + 1.0)
        mdxdy += (ddx * ddy - mdxdy) / (i + 1.0)
    c1 = mdxdy / mdx2
    return my - mx * c1, c1


def _arma_linspace(start, end, num):
    """arma::linspace: start + i (end - start)/(num - 1), the last point set to end."""
    delta = (end - start) / (num - 1.0)
    x = [start + i * delta for i in range(num - 1)]
    return np.array(x + [end])


def _linear_pad(lines, M, axis, xmin, dx, signed_ramp=False):
    """post::FarField::linearPadding (src/farFieldPost.cpp:265-312), line by line:
    real and imaginary parts of entries 0..9 fitted against linspace(0, x_10 - x_0, 10)
    fill the last (M - n)/2 entries with a0 + a1 (x[0] + dxf (i - M + 1)); the same
    entries 0..9 (the reference resets its data pointer to the start) fitted against
    linspace(x_{n-10}, x_n, 10) fill entries n .. n + (M - n)/2 - 1 with
    a0 + a1 (x[9] + (i - n) dxf); dxf is the first grid's step, x_i = xmin + i dx.
    The reference's i - M + 1 mixes an int with Armadillo's unsigned 64-bit n_elem
    (:288), so it is taken modulo 2^64 before the conversion to double (correctly
    rounded, as the conversion is); signed_ramp=True takes the signed difference
    instead (PXG_FIX_LINEAR_PAD)."""
    L2 = np.moveaxis(np.asarray(lines, dtype=np.complex128), axis, 0)
    n = L2.shape[0]
    out = np.zeros((M,) + L2.shape[1:], dtype=np.complex128)
    out[:n] = L2
    added = M - n
    getx = lambda i: xmin + i * dx
    x1 = _arma_linspace(0.0, getx(10) - getx(0), 10)
    x2 = _arma_linspace(getx(n - 10), getx(n), 10)
    dxf = x1[1] - x1[0]
    for j in np.ndindex(*L2.shape[1:]):
        y = L2[(slice(0, 10),) + j]
        c0, c1 = _gsl_fit_linear(x1, y.real)
        d0, d1 = _gsl_fit_linear(x1, y.imag)
        a0, a1 = complex(c0, d0), complex(c1, d1)
        for i in range(M - added // 2, M):
            ramp = float(i - M + 1) if signed_ramp else float((i - M + 1) % (1 << 64))
            out[(i,) + j] = a0 + a1 * (x1[0] + dxf * ramp)
        c0, c1 = _gsl_fit_linear(x2, y.real)
        d0, d1 = _gsl_fit_linear(x2, y.imag)
        a0, a1 = complex(c0, d0), complex(c1, d1)
        for i in range(n, n + added // 2):
            out[(i,) + j] = a0 + a1 * (x2[-1] + (i - n) * dxf)
    return np.moveaxis(out, 0, axis)


def far_field(sol, ref, pad, k, dx, dy, phi_min, phi_max, export_rows=0, export_cols=0,
              padding="zero", xmin=0.0, quirks=0):
    """post::FarField::result for a 3D field (src/farFieldPost.cpp:141-230) with zero
    padding (the default), constant or linear-fit padding (whose fit grids use the
    transverse grid xmin, dx, as the reference's do): subtract the reference field, FFT every column zero-padded
    at its end to the signal length M, fftshift, keep the rows of the y window; then
    the same along the rows of that, keeping the x window; |.|^2 / M; then the
    export resize when export sizes are set (GenericScattering sets 512 x 512).
    The signal length is max(pad, n): the result matrix is taken to start empty.
    sol, ref: (Ny rows, Nx cols) complex arrays; ref may be None."""
    s = np.array(sol, dtype=np.complex128)
    if ref is not None:
        s = s - np.asarray(ref, dtype=np.complex128)
    n = max(s.shape)
    M, (ylo, yhi), (xlo, xhi) = far_field_window(n, pad, k, dx, dy, phi_min, phi_max)
    if yhi < ylo or xhi < xlo:
        raise ValueError("angle window is empty")
    if padding not in ("zero", "constant", "linear"):
        raise ValueError(padding)
    if padding == "constant":
        cols = _constant_pad(s, M, 0)
    elif padding == "linear":
        cols = _linear_pad(s, M, 0, xmin, dx, signed_ramp=bool(quirks & FIX_LINEAR_PAD))
    else:
        cols = s
    ft = _ref_fftshift(np.fft.fft(cols, n=M, axis=0), 0)[ylo:yhi + 1, :]
    # the row pass copies temporary.row(i).t(): Armadillo's t() conjugates complex
    # values (src/farFieldPost.cpp:227), which mirrors the x axis of the pattern;
    # padSignal then pads that conjugated row (:228)
    rows = np.conj(ft)
    if padding == "constant":
        rows = _constant_pad(rows, M, 1)
    elif padding == "linear":
        rows = _linear_pad(rows, M, 1, xmin, dx, signed_ramp=bool(quirks & FIX_LINEAR_PAD))
    ft = _ref_fftshift(np.fft.fft(rows, n=M, axis=1), 1)[:, xlo:xhi + 1]
    res = np.asfortranarray(np.abs(ft) ** 2 / M)
    if export_rows and export_cols:
        res = resize_matrix(res, export_rows, export_cols)
    return res


# ---------------------------------------------------------------------- transmittivity
def closest_index(x, vmin, step):
    """ParaxialSimulation::closestIndex (src/paraxialSimulation.cpp:215-219): the
    truncating conversion of (x - min)/step to unsigned (undefined below 0 in the
    reference; rejected here)."""
    v = (x - vmin) / step
    if not (v >= 0.0 and np.isfinite(v)):
        raise ValueError(f"position {x} below the grid")
    return int(v)


def trapezoid_intensity(col, s, e, xmin, dx):
    """post::Transmittivity::trapezoidalIntegrateIntensityZ (src/transmittivity.cpp:30-42),
    in its order: the two ends, then 2|u|^2 for s < ix < e, times 0.5 (x_e - x_s)/(e - s)
    with the unsigned difference e - s."""
    integral = abs(col[s]) ** 2 + abs(col[e]) ** 2
    for ix in range(s + 1, e):
        integral += 2.0 * abs(col[ix]) ** 2
    span = (e - s) % (1 << 32)
    dxe = ((xmin + e * dx) - (xmin + s * dx)) / float(span) if span else float("nan")
    return integral * dxe * 0.5


def transmittivity_curved(solution, xmin, dx, zmin, dz, R, width):
    """post::Transmittivity::compute after every step n = 1 .. nz-1 of
    CurvedWaveGuideFD::solve (src/transmittivity.cpp:44-83, src/curvedWaveGuide2D.cpp:61-69,
    207-219), no border tracker.  The normalisation is computed on the first call, i.e.
    from the field after step 1, over the guide's z = 0 extent [0, width].
    solution: (nx, nz) complex, column n = field after step n.  Returns (T, I0)."""
    nz = solution.shape[1]
    i0 = trapezoid_intensity(solution[:, 1], closest_index(0.0, xmin, dx),
                             closest_index(width, xmin, dx), xmin, dx)
    out = np.empty(nz - 1)
    for n in range(1, nz):
        z = zmin + n * dz
        xs = -0.5 * z * z / R
        xe = xs + width
        s, e = closest_index(xs, xmin, dx), closest_index(xe, xmin, dx)
        out[n - 1] = trapezoid_intensity(solution[:, n], s, e, xmin, dx) / i0
    return out, i0
