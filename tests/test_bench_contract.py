"""CPU tests of bench.py's contract (no device work): the one-GPU refusal, the workload's
definition and the step-roofline arithmetic."""
from __future__ import annotations

import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert msg["n_gpus"] == 2 and "one GPU" in msg["error"]


def test_ensemble_is_the_seeded_mixed_size_workload():
    sizes = bench.ensemble_sizes(20261202, 512)
    assert sizes.size == 512 and int(sizes.sum()) == 7_482_934
    assert sizes.min() >= 1024 and sizes.max() <= 65536
    assert np.array_equal(sizes, bench.ensemble_sizes(20261202, 512))


def test_step_roofline_arithmetic():
    # 100 B/particle of HBM phases over 1e6 particles = 1e8 B -> 12.5 us at 8 TB/s;
    # 157.3e9 walk flops -> 1 ms at 157.3 TFLOP/s
    r = bench.step_roofline({"a": 60, "b": 40}, 10**6, 157.3e9, 2.0)
    assert abs(r["roofline_ms"] - 1.0125) < 1e-9
    assert abs(r["steps_per_s"] - 500.0) < 1e-9
    assert abs(r["frac"] - 1.0125 / 2.0) < 1e-9
    assert abs(r["roofline_steps_per_s"] - 1e3 / 1.0125) < 1e-6
