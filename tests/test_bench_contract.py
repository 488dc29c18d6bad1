"""CPU checks of bench.py's command-line contract and its work counts."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_work_counts():
    # canonical complex zgeev with vectors: 4 x 25 n^3 real flop
    assert bench.eig_flops(450) == 100.0 * 450 ** 3
    # every term of the solve count is cubic in the harmonics: doubling h
    # multiplies it by 8 exactly, and it grows with each patterned layer
    for n_pat, n_layers in ((1, 3), (4, 6), (6, 8)):
        assert bench.solve_flops(450, n_pat, n_layers) == 8 * bench.solve_flops(225, n_pat,
                                                                                n_layers)
        assert bench.solve_flops(225, n_pat + 1, n_layers) > bench.solve_flops(225, n_pat,
                                                                                n_layers)
    # a solve includes the eigenproblems of its patterned layers (n = 2h)
    assert bench.solve_flops(225, 4, 6) > 4 * bench.eig_flops(450)
    assert bench.FP64_PEAK_TFLOPS == 78.6
    assert bench.FP32_MFMA_PEAK_TFLOPS == 157.3
    assert 1.0e9 < bench.EIG_TRAFFIC_PER_MATRIX_N450 < 2.0e9
