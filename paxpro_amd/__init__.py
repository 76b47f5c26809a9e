"""MI355X-native propagation hot path of PaxPro.

The product is the C ABI library ``paxpro_amd/libpaxgpu.so`` (HIP kernels for
gfx950, declared in ``include/paxgpu.h``) and the C++ drop-in classes under
``paxpro_amd/cpp``.  This Python package is a thin ctypes binding used by the
tests and the benchmark; it has no compute path of its own and raises if the
HIP library is missing.
"""
from .paxgpu import *  # noqa: F401,F403
