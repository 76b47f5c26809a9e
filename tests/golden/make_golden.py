"""Writes the golden fixtures under tests/golden/.

* fft_golden.npz     -- seeded complex inputs and their DFTs from numpy's
                        pocketfft (a textbook DFT library), FFTW sign
                        conventions: forward = numpy.fft.fft, backward =
                        n * numpy.fft.ifft.
* thomas_golden.npz  -- tridiagonal systems solved by the reference's own
                        ThomasAlgorithm, compiled from include/thomasAlgorithm.tpp
                        into oracle/_ref/libref_thomas.so by `make -C oracle ref`.
* csg_kat.json       -- the reference's shape-transform known answers
                        (unittest/transformTest.cpp:4-65) as data.

Run from the repo root:  python tests/golden/make_golden.py
"""
import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def fft_golden():
    rng = np.random.default_rng(0x5EED0F)
    out = {}
    for n in (1, 2, 3, 5, 8, 16, 64, 512):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"x1d_{n}"] = x
        out[f"fwd1d_{n}"] = np.fft.fft(x)
        out[f"bwd1d_{n}"] = np.fft.ifft(x) * n
    for n in (4, 16, 64):
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        out[f"x2d_{n}"] = a
        out[f"fwd2d_{n}"] = np.fft.fft2(a)
        out[f"bwd2d_{n}"] = np.fft.ifft2(a) * n * n
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)


def thomas_golden():
    path = os.path.join(ROOT, "oracle", "_ref", "libref_thomas.so")
    lib = C.CDLL(path)
    lib.ref_thomas_c128.argtypes = [C.c_void_p] * 3 + [C.c_uint]
    rng = np.random.default_rng(0x5EED07)
    out = {}
    cases = []
    for n in (2, 3, 7, 64, 257, 1024):
        for kind in ("dominant", "cn"):
            if kind == "dominant":
                sub = rng.standard_normal(n - 1) + 1j * rng.standard_normal(n - 1)
                diag = 4.0 + rng.standard_normal(n) + 1j * rng.standard_normal(n)
            else:
                rho = 8.328
                diag = np.full(n, 1 + 0.5j * rho) + 1e-3 * rng.standard_normal(n)
                sub = np.full(n - 1, -0.25j * rho)
            rhs = rng.standard_normal(n) + 1j * rng.standard_normal(n)
            d, s, r = diag.copy(), sub.copy(), rhs.copy()
            lib.ref_thomas_c128(d.ctypes.data, s.ctypes.data, r.ctypes.data, n)
            key = f"{kind}_{n}"
            out[f"diag_{key}"], out[f"sub_{key}"], out[f"rhs_{key}"] = diag, sub, rhs
            out[f"x_{key}"] = d
            cases.append(key)
    out["cases"] = np.array(cases)
    np.savez_compressed(os.path.join(HERE, "thomas_golden.npz"), **out)


def csg_kat():
    kats = [
        {"name": "rotateZ", "op": ["rotate", 90.0, 2], "p": [1, 0, 0], "expect": [0, -1, 0]},
        {"name": "rotateY", "op": ["rotate", 90.0, 1], "p": [1, 0, 0], "expect": [0, 0, 1]},
        {"name": "rotateX", "op": ["rotate", 90.0, 0], "p": [0, 1, 0], "expect": [0, 0, -1]},
        {"name": "translation", "op": ["translate", 10.0, -10.0, -3.0], "p": [1, -1, 10],
         "expect": [-9, 9, 13]},
    ]
    with open(os.path.join(HERE, "csg_kat.json"), "w") as f:
        json.dump({"source": "unittest/transformTest.cpp:4-65", "tol": 1e-6, "kats": kats}, f,
                  indent=1)


if __name__ == "__main__":
    fft_golden()
    thomas_golden()
    csg_kat()
