"""Helpers shared by the GPU parity tests."""
import sys

import numpy as np

import paxpro_amd as P
import pyoracle as O


def announce(*args):
    print("[gpu-case]", *args, flush=True)
    sys.stdout.flush()


def rel_l2(a, ref):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(ref)) / np.linalg.norm(np.asarray(ref)))


def rel_l2_int(a, ref):
    ia, ir = np.abs(a) ** 2, np.abs(ref) ** 2
    return float(np.linalg.norm(ia - ir) / np.linalg.norm(ir))


def fft3d_problem(g, n, nz, precision=P.PREC_FP64, batch=1, absorber=None, store=(0, 0), quirks=0):
    ab = (0, 0, 0.0, 0.0) if absorber is None else (absorber.npix_x, absorber.npix_y,
                                                     absorber.inv_x, absorber.inv_y)
    return P.problem(P.FFT3D, n, nz, g.dx, g.dz, g.k, xmin=g.xmin, ymin=g.ymin, dy=g.dy,
                     zmin=g.zmin, precision=precision, batch=batch, absorb=ab,
                     store_dsx=store[0], store_dsz=store[1], quirks=quirks)
