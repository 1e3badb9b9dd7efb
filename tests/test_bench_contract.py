"""bench.py's contract on the CPU: the one-GPU guard and the step byte model."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout.strip() == ""


def test_step_byte_model_of_the_multispecies_workload():
    import bench
    from autoverse_amd import workloads
    cfg, _ = workloads.multispecies()
    assert (cfg.width, cfg.height, cfg.channels, len(cfg.kernels)) == (1024, 1024, 3, 9)
    # 3 source channels, 9 kernels, S = 8 * 513 * 1024, N = 1024^2:
    # 3 (4N + 3S) + 4 * 9 S + 4 * 3 N + 12 * 3 N
    n, s = 1024 * 1024, 8 * 513 * 1024
    assert bench.step_model_bytes(cfg) == 3 * (4 * n + 3 * s) + 36 * s + 12 * n + 36 * n == 252026880


@pytest.mark.parametrize("dims", [(64, 64, 1), (32, 32, 32)])
def test_step_byte_model_counts_depth(dims):
    import bench
    from autoverse_amd import workloads
    cfg, _ = workloads.multispecies()
    w, h, d = dims
    cfg.width, cfg.height, cfg.depth = w, h, d
    n, s = w * h * d, 8 * (w // 2 + 1) * h * d
    cs = len({k.source_channel for k in cfg.kernels})
    assert bench.step_model_bytes(cfg) == cs * (4 * n + 3 * s) + 4 * len(cfg.kernels) * s + 16 * cfg.channels * n
