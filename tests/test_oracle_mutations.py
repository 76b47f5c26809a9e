"""Mutation check of the oracle's closed-form pins (CPU only).

Each case below plants one plausible mistake in a copy of oracle/paxoracle.c
(a doubled diffraction coefficient, a flipped sign, swapped axes, a doubled
CN rho, a doubled ADI x-diffusion), builds the copy into a temporary
directory, and runs tests/test_oracle_closed_forms.py against it through
PAXORACLE_LIB.  Every mutant must fail that file; the unmutated copy, built
the same way, must pass it.  The mistakes live only in the temporary copy:
nothing here touches the GPU or the oracle the other tests load.

Each replacement names the exact text it changes and how often it occurs, so
the test fails loudly (instead of silently testing nothing) if the oracle's
source drifts away from these lines.
"""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(ROOT, "oracle", "paxoracle.c")
HDR = os.path.join(ROOT, "oracle", "paxoracle.h")
CLOSED_FORMS = os.path.join(HERE, "test_oracle_closed_forms.py")

# (old text, new text, expected occurrences of the old text)
_ADI_X = [
    ("diag[indx] = 1.0 / dz + I / (k * dx * dx) + beta * k + I * delta * k;",
     "diag[indx] = 1.0 / dz + I / (0.5 * k * dx * dx) + beta * k + I * delta * k;", 1),
    ("if (j > 0) sub[indx - 1] = -0.5 * I / (k * dx * dx);",
     "if (j > 0) sub[indx - 1] = -0.5 * I / (0.5 * k * dx * dx);", 1),
    ("diag[idx] -= 0.5 * I * cexp(I * kd) / (k * dx * dx);",
     "diag[idx] -= 0.5 * I * cexp(I * kd) / (0.5 * k * dx * dx);", 2),
]

MUTATIONS = {
    "fft3d_coefficient_doubled": [
        ("const orc_c A = I * (0.5 * dz / k);", "const orc_c A = I * (1.0 * dz / k);", 1)],
    "fft3d_kernel_sign_flipped": [
        ("cur[i] *= cexp(-A * (kx * kx + ky * ky));", "cur[i] *= cexp(A * (kx * kx + ky * ky));", 1)],
    "fft3d_kx_from_row_ky_from_column": [
        ("double kx = kxs[col], ky = kys[row];", "double kx = kxs[row], ky = kys[col];", 1)],
    "fft3d_kx_with_dy": [
        ("kxs[i] = spatial_freq(i, n, g->dx);", "kxs[i] = spatial_freq(i, n, g->dy);", 1)],
    "fft2d_coefficient_doubled": [
        ("const orc_c A = I * (0.5 * g->dz / g->k);", "const orc_c A = I * (1.0 * g->dz / g->k);", 1)],
    "cn_rho_doubled": [
        ("const double rho = g->dz / (g->k * g->dx * g->dx);",
         "const double rho = 2.0 * g->dz / (g->k * g->dx * g->dx);", 1)],
    "adi_x_diffusion_doubled_with_boundary_terms": _ADI_X,
}


def _build(tmp, text):
    src = os.path.join(tmp, "paxoracle.c")
    with open(src, "w") as f:
        f.write(text)
    shutil.copy(HDR, os.path.join(tmp, "paxoracle.h"))
    lib = os.path.join(tmp, "liboracle.so")
    subprocess.run(["gcc", "-O2", "-fopenmp", "-fPIC", "-std=gnu99", "-shared", src, "-o", lib, "-lm"],
                   check=True, capture_output=True)
    return lib


def _closed_forms(lib):
    env = dict(os.environ, PAXORACLE_LIB=lib, OMP_NUM_THREADS="2")
    env["PYTHONPATH"] = os.pathsep.join([os.path.join(ROOT, "oracle"), HERE, env.get("PYTHONPATH", "")])
    return subprocess.run([sys.executable, "-m", "pytest", CLOSED_FORMS, "-x", "-q", "-p", "no:cacheprovider",
                           "-m", "not gpu"], cwd=ROOT, env=env, capture_output=True, text=True)


def _source():
    with open(SRC) as f:
        return f.read()


def test_unmutated_copy_passes(tmp_path):
    r = _closed_forms(_build(str(tmp_path), _source()))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_mutant_fails_closed_forms(name, tmp_path):
    text = _source()
    for old, new, count in MUTATIONS[name]:
        assert text.count(old) == count, f"{name}: expected {count} of {old!r} in paxoracle.c"
        text = text.replace(old, new)
    r = _closed_forms(_build(str(tmp_path), text))
    assert r.returncode == 1, f"{name} was not caught:\n" + r.stdout[-2000:] + r.stderr[-2000:]
