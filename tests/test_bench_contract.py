"""bench.py contract checks that need no GPU: more than one GPU is refused
with a message, and the committed rocprofv3 counter summary covers every
work-group shape of both FFT3D passes that the autotuner can pick."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_traffic_summary_covers_both_passes():
    B = _bench()
    for kernel in ("pass_x", "pass_y"):
        for lines in (1, 2, 4, 8, 16):
            t = B.measured_traffic(kernel, 64, lines)
            assert t is not None, (kernel, lines)
            alg = 2 * 8 * 512 * 512 * 64
            assert alg <= t < 1.1 * alg, (kernel, lines, t)
    assert B.measured_traffic("pass_x", 32, 8) is None
