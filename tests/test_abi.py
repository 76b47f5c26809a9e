"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/paxgpu.h declares, and rejects bad input without touching a GPU."""
import ctypes as C
import os
import re

import pytest

import paxpro_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "paxgpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char \*)\s*(pxg_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = P.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(P.EXPORTS)


def test_version_string():
    assert "gfx950" in P.version()


def test_plan_bytes_and_validation_without_gpu():
    prob = P.problem(P.FFT3D, 512, 1001, 39.0625, 10.0, 62.83, batch=4, precision=P.PREC_MIXED)
    b = P.plan_bytes(prob)
    assert b >= 512 * 512 * 4 * 8
    bad = P.problem(P.FFT3D, 5000, 1001, 4.0, 10.0, 62.83)   # beyond 4096 points per side
    with pytest.raises(P.PxgError) as e:
        P.plan_bytes(bad)
    assert e.value.code == P.ERR_UNSUPPORTED
    bad2 = P.problem(P.CN, 1, 10, 1.0, 1.0, 1.0)
    with pytest.raises(P.PxgError) as e:
        P.plan_bytes(bad2)
    assert e.value.code == P.ERR_INVALID
    ds = P.problem(P.CN, 64, 10, 1.0, 1.0, 1.0, store_dsx=2, store_dsz=1)
    with pytest.raises(P.PxgError) as e:
        P.plan_bytes(ds)
    assert e.value.code == P.ERR_UNSUPPORTED


def test_null_arguments_are_rejected():
    L = P.lib()
    assert L.pxg_plan_create(None, None, None) == P.ERR_INVALID
    assert L.pxg_step(None, 1, 1) == P.ERR_INVALID
    assert b"null" in L.pxg_last_error()


def test_plan_bytes_count_the_material_maps():
    """FFT3D plans reserve one material map (2 + 16 B per cell) for batches whose
    map pass X builds (8 to 15 fields), two for batches marched in halves (16 or
    more, even), and a ring of 8 for fewer than 8 fields (maps of 8 steps per
    launch); lengths without a power-of-two kernel keep one."""
    def b(n, batch):
        return P.plan_bytes(P.problem(P.FFT3D, n, 101, 20000.0 / n, 10.0, 62.83, batch=batch,
                                      precision=P.PREC_MIXED))
    for n in (256, 512):
        per_map = n * n * (2 + 16)
        per_field = b(n, 9) - b(n, 8)
        assert per_field > 0
        assert b(n, 16) - b(n, 8) == 8 * per_field + per_map
        assert b(n, 7) - b(n, 8) == -per_field + 7 * per_map
        assert b(n, 1) - b(n, 8) == -7 * per_field + 7 * per_map
    per_field_511 = b(511, 9) - b(511, 8)
    assert b(511, 7) - b(511, 8) == -per_field_511
