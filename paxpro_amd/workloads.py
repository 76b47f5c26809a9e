"""Synthetic workloads of the propagation hot path, as plain numpy inputs for
the C ABI (no oracle involved): grids, CSG phantoms, waveguide job lists and
the served request trace.

Each generator states the workload it builds; the random ones draw from a
SplitMix64 stream with a fixed seed, so every run sees the same inputs.
"""
import math

import numpy as np

from . import paxgpu as P

LAMBDA = 0.1                      # nm
K_01 = 2 * math.pi / LAMBDA       # nm^-1


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


def spheres_csg(spheres):
    """CSG arrays of spheres given as (cx, cy, cz, r, delta, beta): each sphere
    is its own UNION part, all parts in one module (the last one wins where
    spheres overlap).  A shape's transform maps world points to the shape
    frame; a translation by c is the identity with -c in the last column, the
    form geom::Shape::translate builds (src/shape.cpp:16-40)."""
    n = len(spheres)
    shapes = np.zeros(n, P.SHAPE_DTYPE)
    parts = np.zeros(n, P.PART_DTYPE)
    for i, (cx, cy, cz, r, d, b) in enumerate(spheres):
        T = np.eye(4)
        T[0, 3], T[1, 3], T[2, 3] = -cx, -cy, -cz
        shapes[i]["type"], shapes[i]["op"] = P.SHAPE_SPHERE, P.OP_UNION
        shapes[i]["T"], shapes[i]["p"] = T.ravel(), (r, 0.0, 0.0)
        parts[i]["first_shape"], parts[i]["n_shapes"], parts[i]["op"] = i, 1, P.OP_UNION
        parts[i]["delta"], parts[i]["beta"] = d, b
    modules = np.zeros(1, P.MODULE_DTYPE)
    modules[0]["first_part"], modules[0]["n_parts"] = 0, n
    return shapes, parts, modules


def random_spheres(seed, count, xy_range, z_range, r_range, d_range, beta_ratio=50.0):
    """`count` spheres: radius, centre x, y, z and delta drawn uniformly in that
    order per sphere; beta = delta / beta_ratio."""
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


def gaussian_field(n, dx, xmin, waist):
    """GaussianBeam at its waist, centred, no tilt: exp(-(x^2 + y^2)/w^2) on
    an n x n grid, column-major with rows = y (src/gaussianBeam.cpp:34-63,
    src/paraxialSimulation.cpp:381-395)."""
    x = xmin + dx * np.arange(n)
    gx = np.exp(-(x / waist) ** 2)
    return np.asfortranarray(np.outer(gx, gx).astype(np.complex128))


# --- 2D Crank-Nicolson slab waveguide (Nx = 1024, 2000 steps) ---------------

def cn_slab(precision=P.PREC_FP64, nsteps=2000, nx=1024, absorb=50):
    """Plane wave (u = 1) into a slab guide: core -50 < x < 50 nm vacuum,
    cladding delta 4.5e-5, beta 3.5e-6; x in [-500, 500] nm with 1024 nodes,
    dz = 500 nm, lambda 0.1 nm, transparent boundaries and a 50-cell absorbing
    layer (1/pixel 0.1).  Returns (problem, material planes, initial field)."""
    xmin, dx, dz = -500.0, 1000.0 / (nx - 1), 500.0
    x = xmin + dx * np.arange(nx)
    inside = (x > -50.0) & (x < 50.0)
    db = np.empty((1, nx, 2))
    db[0, :, 0] = np.where(inside, 0.0, 4.5e-5)
    db[0, :, 1] = np.where(inside, 0.0, 3.5e-6)
    prob = P.problem(P.CN, nx, nsteps + 1, dx, dz, K_01, xmin=xmin, bc=P.BC_TRANSPARENT,
                     absorb=(absorb, 0, 0.1, 0.0), precision=precision)
    return prob, db, np.ones(nx, complex)


# --- 3D ADI, 2048^2 over 40 um, 500 steps, 64 spheres -----------------------

def adi_phantom(seed=0x5EED0002, count=64):
    """64 spheres: radius U[500, 3000], centre x, y U[-17000, 17000], z
    U[-2000, 2000] nm, delta U[1e-6, 1e-5], beta = delta / 50."""
    return random_spheres(seed, count, (-17000.0, 17000.0), (-2000.0, 2000.0), (500.0, 3000.0),
                          (1e-6, 1e-5))


def adi_problem(precision=P.PREC_FP64, n=2048, nsteps=500):
    """dx = 40000/2048 nm, z in [-5000, 5000] nm with dz = 20, lambda 0.1 nm,
    transparent boundaries; source a Gaussian of waist 8000 nm."""
    dx = 40000.0 / n
    xmin = -20000.0
    dz = 10000.0 / nsteps
    prob = P.problem(P.ADI, n, nsteps + 1, dx, dz, K_01, xmin=xmin, zmin=-5000.0, use_tbc=1,
                     precision=precision)
    return prob, gaussian_field(n, dx, xmin, 8000.0)


# --- 256 curved-waveguide CN jobs of uneven Nx -------------------------------

def curved_guide_jobs(seed=0x5EED0003, count=256, nsteps=1000, dz=100.0, nx_range=(256, 8192)):
    """Per job: Nx = round(exp U[ln 256, ln 8192]), lambda U[0.05, 0.2] nm,
    radius R U[1e6, 5e7] nm, width w U[50, 200] nm; length L = nsteps * dz;
    x window [-L^2/(2R) - 2w - 200, w + 200] with Nx nodes; cladding delta
    4.5e-5, beta 3.5e-6."""
    g = splitmix64(seed)
    jobs = []
    L = nsteps * dz
    for _ in range(count):
        nx = int(round(math.exp(math.log(nx_range[0]) + (math.log(nx_range[1]) - math.log(nx_range[0])) * next(g))))
        lam = 0.05 + 0.15 * next(g)
        R = 1e6 + 49e6 * next(g)
        w = 50.0 + 150.0 * next(g)
        lo, hi = -L * L / (2 * R) - 2 * w - 200.0, w + 200.0
        dx = (hi - lo) / (nx - 1)
        jobs.append(dict(nx=nx, lam=lam, R=R, width=w, xmin=lo, dx=dx, dz=dz, nsteps=nsteps,
                         wg=dict(R=R, width=w, clad_delta=4.5e-5, clad_beta=3.5e-6, dx=dx, dz=dz)))
    return jobs


def curved_guide_problem(job, precision):
    return P.problem(P.CN, job["nx"], job["nsteps"] + 1, job["dx"], job["dz"], 2 * math.pi / job["lam"],
                     xmin=job["xmin"], bc=P.BC_TRANSPARENT, precision=precision)


# --- served request trace -----------------------------------------------------

def served_trace(count, rate=200.0, seed=0x5EED0004, precision=P.PREC_MIXED):
    """Poisson arrivals at `rate` per second.  70 %: FFT3D, N uniform in {256,
    512, 1024, 2048} over 20 um, steps uniform in 200..2000 (dz = 10 nm), a
    64-sphere phantom from a per-request seed (the ADI phantom's generator with
    radii and transverse centres halved for the 20 um window); 30 %: one
    curved-waveguide CN job drawn as above.  Returns a list of dicts with the
    arrival time (s), the problem and the material; initial fields are a
    Gaussian of waist 3 um (FFT3D) or a plane wave (CN)."""
    g = splitmix64(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += -math.log(1.0 - next(g)) / rate
        if next(g) < 0.7:
            n = (256, 512, 1024, 2048)[min(3, int(4 * next(g)))]
            nsteps = 200 + min(1800, int(1801 * next(g)))
            dx = 20000.0 / n
            dz = 10.0
            zmin = -nsteps * dz / 2
            prob = P.problem(P.FFT3D, n, nsteps + 1, dx, dz, K_01, xmin=-10000.0, zmin=zmin,
                             precision=precision)
            sph = random_spheres(seed + 1 + i, 64, (-8500.0, 8500.0), (-2000.0, 2000.0), (250.0, 1500.0),
                                 (1e-6, 1e-5))
            out.append(dict(t=t, kind="fft3d", n=n, nsteps=nsteps, prob=prob, csg=spheres_csg(sph),
                            cells=n * n))
        else:
            job = curved_guide_jobs(seed + 1 + i, 1)[0]
            out.append(dict(t=t, kind="cn", n=job["nx"], nsteps=job["nsteps"],
                            prob=curved_guide_problem(job, precision), curved=job["wg"], cells=job["nx"]))
    return out
