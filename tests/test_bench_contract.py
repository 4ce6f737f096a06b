"""CPU tests of bench.py's contract that need no GPU: the one-GPU refusal for every workload, and the
committed rocprofv3 traffic summary that the benchmark line's roofline.traffic is read from."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("workload", ["spin_glass_n64_L5", "batched", "learned", "legacy"])
def test_more_than_one_gpu_is_refused(workload):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--workload", workload],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"] and line["n_gpus"] == 2


def test_traffic_summary_feeds_the_roofline():
    sys.path.insert(0, ROOT)
    import bench
    for precision in ("fp64", "fp32"):
        t = bench._profiled_traffic("spin_glass_n64_L5", precision, "hermitian")
        assert t is not None and t > 0
    # FP64 moves twice the bytes of complex64, and stays below the full-block algorithmic figure
    n, h = 64, 6188
    alg = 14 * n * n * 16 * h / 4
    t64 = bench._profiled_traffic("spin_glass_n64_L5", "fp64", "hermitian")
    t32 = bench._profiled_traffic("spin_glass_n64_L5", "fp32", "hermitian")
    assert 1.7 < t64 / t32 < 2.5
    assert t64 < alg
    assert bench._profiled_traffic("spin_glass_n8_L4", "fp64", "hermitian") is None
    assert bench._profiled_traffic("spin_glass_n64_L5", "fp64", "full") is None
