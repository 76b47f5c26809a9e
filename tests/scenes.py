"""Shared synthetic inputs for parity tests and the benchmark: grids, CSG
phantoms and sources, built with the oracle's shape-transform builders
(pinned by the reference's known answers) into the flat CSG arrays that both
the oracle and libpaxgpu consume."""
import numpy as np

import pyoracle as O

K_01 = 2 * np.pi / 0.1  # lambda = 0.1 nm


def splitmix64(seed):
    """SplitMix64 stream (Steele, Lea, Flood 2014) of floats in [0, 1)."""
    s = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        s = (s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        z ^= z >> 31
        yield (z >> 11) * (1.0 / (1 << 53))


def spheres_scene(spheres):
    """spheres: list of (cx, cy, cz, r, delta, beta); each sphere is its own
    part, all parts in one module (the last added wins on overlap)."""
    n = len(spheres)
    shapes = np.zeros(n, O.SHAPE_DTYPE)
    parts = np.zeros(n, O.PART_DTYPE)
    for i, (cx, cy, cz, r, d, b) in enumerate(spheres):
        T = O.mat4_identity()
        O.translate(T, cx, cy, cz)
        shapes[i]["type"], shapes[i]["op"], shapes[i]["T"], shapes[i]["p"] = O.SPHERE, O.UNION, T, (r, 0, 0)
        parts[i]["first_shape"], parts[i]["n_shapes"], parts[i]["op"] = i, 1, O.UNION
        parts[i]["delta"], parts[i]["beta"] = d, b
    modules = np.zeros(1, O.MODULE_DTYPE)
    modules[0]["first_part"], modules[0]["n_parts"] = 0, n
    return shapes, parts, modules


def random_spheres(seed, count, xy_range, z_range, r_range, d_range, beta_ratio=50.0):
    g = splitmix64(seed)
    out = []
    for _ in range(count):
        r = r_range[0] + (r_range[1] - r_range[0]) * next(g)
        cx = xy_range[0] + (xy_range[1] - xy_range[0]) * next(g)
        cy = xy_range[0] + (xy_range[1] - xy_range[0]) * next(g)
        cz = z_range[0] + (z_range[1] - z_range[0]) * next(g)
        d = d_range[0] + (d_range[1] - d_range[0]) * next(g)
        out.append((cx, cy, cz, r, d, d / beta_ratio))
    return out


def fft3d_case(n, dz=10.0, nsteps=1000, k=K_01, extent=20000.0):
    """cfg-1 style grid scaled to n cells over `extent` nm, z in [-nsteps*dz/2, ...]."""
    dx = extent / n
    xmin = -extent / 2
    zmin = -nsteps * dz / 2
    g = O.grid(xmin, xmin + (n - 1) * dx, dx, zmin, zmin + nsteps * dz, dz, k)
    return g


def gaussian_field(g, n, waist, k=K_01):
    gb = O.Gauss(k, 1.0, 0.0, waist, 0.0, 0.0, 0.0, 0.0)
    return O.gauss3d_field(g, gb, n)


X_AXIS, Y_AXIS, Z_AXIS = 0, 1, 2


def _shape(kind, p, op=None):
    s = np.zeros(1, O.SHAPE_DTYPE)[0]
    s["type"], s["op"], s["T"], s["p"] = kind, O.UNION if op is None else op, O.mat4_identity(), p
    return s


def primitive_geom_scene(s=100.0, delta=8e-6, beta=2e-7):
    """Examples/primitiveGeomExample.cpp:15-33 with every length times s (nm):
    one part holding a sphere, a box rotated 45 deg about z and moved along x,
    a cylinder carved out of them (DIFFERENCE); the part is then rotated 30 deg
    about x and moved along x (Part::rotate/translate act on the shapes added so
    far, src/geometry.cpp:153-175), and a sphere scaled 4x along x (an
    ellipsoid) is added after the part's transforms."""
    sphere = _shape(O.SPHERE, (10 * s, 0, 0))
    box = _shape(O.BOX, (20 * s, 20 * s, 20 * s))
    O.rotate(box["T"], 45.0, Z_AXIS)
    O.translate(box["T"], 5 * s, 0.0, 0.0)
    cyl = _shape(O.CYLINDER, (10 * s, 10 * s, 10 * s), op=O.DIFFERENCE)
    O.translate(cyl["T"], 15 * s, 0.0, 0.0)
    for sh in (sphere, box, cyl):
        O.rotate(sh["T"], 30.0, X_AXIS)
        O.translate(sh["T"], 20 * s, 0.0, 0.0)
    ell = _shape(O.SPHERE, (10 * s, 0, 0))
    O.scale(ell["T"], 4.0, X_AXIS)
    O.translate(ell["T"], -20 * s, 0.0, 0.0)
    shapes = np.array([sphere, box, cyl, ell], dtype=O.SHAPE_DTYPE)
    parts = np.zeros(1, O.PART_DTYPE)
    parts[0] = (0, 4, O.UNION, 0, delta, beta)
    modules = np.zeros(1, O.MODULE_DTYPE)
    modules[0] = (0, 1)
    return shapes, parts, modules


def slab_example_scene():
    """Examples/slab.cpp:14-27: a 250 x 50 x 5000 nm box slab (delta 1e-4,
    beta 1e-6) moved to y = -36.5 nm and tilted -0.2 deg about x, one module."""
    box = _shape(O.BOX, (250.0, 50.0, 5000.0))
    O.translate(box["T"], 0.0, -36.5, 0.0)
    O.rotate(box["T"], -0.2, X_AXIS)
    shapes = np.array([box], dtype=O.SHAPE_DTYPE)
    parts = np.zeros(1, O.PART_DTYPE)
    parts[0] = (0, 1, O.UNION, 0, 1e-4, 1e-6)
    modules = np.zeros(1, O.MODULE_DTYPE)
    modules[0] = (0, 1)
    return shapes, parts, modules


def multi_module_scene(s=1000.0):
    """Two modules.  Module 0: a box part, a rotated cylinder part and a sphere
    part carved out of the module (Module::difference, src/geometry.cpp:194-233:
    the DIFFERENCE part claims its points with delta = beta = 0).  Module 1: a
    larger sphere made of a cone (cylinder with r1 != r2) minus a scaled sphere,
    which only sets the material where module 0 does not claim the point."""
    box = _shape(O.BOX, (4 * s, 3 * s, 2 * s))
    O.rotate(box["T"], 20.0, Z_AXIS)
    cyl = _shape(O.CYLINDER, (1.2 * s, 1.2 * s, 6 * s))
    O.rotate(cyl["T"], 90.0, Y_AXIS)
    O.translate(cyl["T"], 0.5 * s, 1.0 * s, 0.0)
    hole = _shape(O.SPHERE, (1.1 * s, 0, 0))
    O.translate(hole["T"], -1.0 * s, 0.0, 0.3 * s)
    cone = _shape(O.CYLINDER, (3.5 * s, 1.5 * s, 5 * s))
    O.rotate(cone["T"], 35.0, X_AXIS)
    carve = _shape(O.SPHERE, (1.0 * s, 0, 0), op=O.DIFFERENCE)
    O.scale(carve["T"], 2.0, Y_AXIS)
    O.translate(carve["T"], 2.5 * s, -1.5 * s, 0.0)
    shapes = np.array([box, cyl, hole, cone, carve], dtype=O.SHAPE_DTYPE)
    parts = np.zeros(4, O.PART_DTYPE)
    parts[0] = (0, 1, O.UNION, 0, 6e-6, 1.2e-7)
    parts[1] = (1, 1, O.UNION, 0, 9e-6, 3e-7)
    parts[2] = (2, 1, O.DIFFERENCE, 0, 5e-5, 5e-5)   # delta/beta unused: carved to 0
    parts[3] = (3, 2, O.UNION, 0, 3e-6, 6e-8)
    modules = np.zeros(2, O.MODULE_DTYPE)
    modules[0], modules[1] = (0, 3), (3, 1)
    return shapes, parts, modules
