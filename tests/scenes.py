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
