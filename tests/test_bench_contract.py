"""bench.py's contract on the CPU: more than one GPU is refused before any
GPU call, and the per-H.psi byte model it reports against reproduces the
figures the workload definition states (1.21 / 0.62 MB for Si8 and
9.68 / 4.96 MB for Si64, FP64 / FP32)."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr or "single" in r.stderr
    assert r.stdout.strip() == ""


@pytest.mark.parametrize("name,rep,n,mb64,mb32", [("si8", 1, 32, 1.21, 0.62),
                                                  ("si64", 2, 64, 9.68, 4.96)])
def test_survey_byte_model(name, rep, n, mb64, mb32):
    sys.path.insert(0, REPO)
    import bench
    from sdft_amd import systems
    s = systems.build_system(name, 10.26, rep, n)
    assert abs(bench.survey_bytes_per_hpsi(s, 16, 8) / 1e6 - mb64) < 0.005
    assert abs(bench.survey_bytes_per_hpsi(s, 8, 4) / 1e6 - mb32) < 0.005


def test_plane_kernel_bytes_si64():
    """The dominant kernel's algorithmic bytes per launch at Si64 FP32:
    W read and written once (2 x 128 pairs x 64 planes x 401 columns x 8 B)
    plus v_KS once (64^3 x 4 B) = 53.6 MB."""
    sys.path.insert(0, REPO)
    import bench
    from sdft_amd import systems
    s = systems.build_system("si64", 10.26, 2, 64)
    ncol = bench.ctx_ncol(s)
    assert ncol == 401
    b, _ = bench.plane_kernel_model(s, ncol, 128, 8, 4, 31)
    assert b == 2 * 128 * 64 * 401 * 8 + 64 ** 3 * 4
    assert abs(b / 1e6 - 53.6) < 0.05
