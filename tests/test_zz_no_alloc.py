"""pxg_thomas_batched_noalloc, the entry point without device allocation (no
reference counterpart; for callers that cannot allocate, e.g. under graph
capture): on device arrays it gives pxg_thomas_batched's results bit for bit,
on two calls in a row (the second smaller, so stale scratch from the first
cannot hide), for both layouts, and it rejects sizes beyond its scratch before
launching anything.  Device buffers come from the HIP runtime that libpaxgpu.so
itself links."""
import ctypes as C

import numpy as np
import pytest

import paxpro_amd as P
import pyoracle as O
from gpu_common import announce

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2   # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    P.lib()  # loads the library and, with it, its HIP runtime
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


class Dev:
    """A device copy of a complex128 numpy array."""

    def __init__(self, hip, a):
        self.hip, self.a = hip, np.ascontiguousarray(a, dtype=np.complex128)
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(16, self.a.nbytes)) == 0
        assert hip.hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _systems(rng, nsys, n, layout):
    shape = (nsys, n) if layout == P.LAYOUT_CONTIGUOUS else (n, nsys)
    sshape = (nsys, n - 1) if layout == P.LAYOUT_CONTIGUOUS else (n - 1, nsys)
    cplx = lambda sh: rng.standard_normal(sh) + 1j * rng.standard_normal(sh)
    return 4.0 + cplx(shape), cplx(sshape), cplx(shape)


@pytest.fixture(scope="module")
def ctx():
    c = P.Ctx(0)
    yield c
    c.close()


@pytest.mark.parametrize("layout", [P.LAYOUT_CONTIGUOUS, P.LAYOUT_INTERLEAVED])
def test_noalloc_thomas_equals_allocating_entry_point(ctx, layout):
    hip = _hip()
    rng = np.random.default_rng(20 + layout)
    for nsys, n in ((700, 1300), (37, 129)):   # 910,000 rows, then a smaller call
        announce("thomas noalloc", layout, nsys, n)
        d, s, r = _systems(rng, nsys, n, layout)
        want_x, want_r = ctx.thomas_batched(d, s, r, layout)
        dd, ds, dr = Dev(hip, d), Dev(hip, s), Dev(hip, r)
        ctx.thomas_batched_noalloc(dd.p.value, ds.p.value, dr.p.value, nsys, n, layout)
        assert hip.hipDeviceSynchronize() == 0
        got_x, got_r = dd.get(), dr.get()
        for a in (dd, ds, dr):
            a.free()
        np.testing.assert_array_equal(got_x, want_x)
        np.testing.assert_array_equal(got_r, want_r)
        # and one system against the oracle's Thomas (the reference's recurrence)
        if layout == P.LAYOUT_CONTIGUOUS:
            ref = O.thomas(d[3], s[3], r[3])
            assert np.max(np.abs(got_x[3] - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_noalloc_thomas_rejects_sizes_beyond_its_scratch(ctx):
    hip = _hip()
    nsys, n = 2, P.THOMAS_NOALLOC_MAX_ROWS // 2 + 1
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), 16) == 0   # never touched: the call must fail first
    with pytest.raises(P.PxgError) as e:
        ctx.thomas_batched_noalloc(p.value, p.value, p.value, nsys, n)
    hip.hipFree(p)
    assert e.value.code == P.ERR_UNSUPPORTED and "scratch" in str(e.value)
    with pytest.raises(P.PxgError) as e:
        ctx.thomas_batched_noalloc(None, None, None, 4, 4)
    assert e.value.code == P.ERR_INVALID
