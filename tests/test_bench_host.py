"""CPU tests of the measurement harness (bench.py): workload sizes, flop
counts and the one-GPU guard -- no device work."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_headline_workloads():
    # 48^4 'abcd,dbef->acef': m = n = k = 2304, 8 m n k real flops per complex MAC
    fl, (m, n, k) = bench.flops(bench.CONFIGS["c2"])
    assert (m, n, k) == (2304, 2304, 2304) and fl == 8 * 2304 ** 3
    assert bench.CONFIGS["c2"]["dtype"] == "complex64" and bench.CONFIGS["c2"]["peak"] == "fp32"
    # 256x4x256 'asb,btc->astc': m = n = 1024, k = 256
    fl, (m, n, k) = bench.flops(bench.CONFIGS["c3"])
    assert (m, n, k) == (1024, 1024, 256) and fl == 8 * 1024 * 1024 * 256
    assert bench.CONFIGS["c3"]["dtype"] == "complex128" and bench.CONFIGS["c3"]["peak"] == "fp64"


def test_peaks_are_dense_mi355x_figures():
    assert bench.PEAK["fp32"] == 157.3 and bench.PEAK["fp64"] == 78.6


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
