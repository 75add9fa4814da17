"""bench.py's host logic on the CPU: the one-GPU contract, the committed
counter-profile figures it reads, and the host CPU description of its CPU
baseline."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_transport_traffic_per_history_is_read_from_the_profile():
    t = bench.measured_traffic_per_history()
    assert t is not None and 50.0 < t < 5000.0
    # every profiled workload has its own figure; the C5 grid misses the cache
    # far more often than C2's, the uniform pencil cube reads no voxel at all
    per = {w: bench.measured_traffic_per_history(w) for w in bench.TRAFFIC_PROFILES}
    assert per["ct-spiral-256"] == t
    assert per["ct-spiral-512-dual"] > 2 * t
    assert 0.0 < per["pencil-64"] < t
    assert bench.measured_traffic_per_history("ct-segmentation") is None


def test_host_cpu_description():
    h = bench.host_cpu()
    assert h["logical_cpus"] >= 1 and 1 <= h["allowed_cpus"] <= h["logical_cpus"]
    assert 1 <= h["threads"] <= h["allowed_cpus"]
    if h["cgroup_cpu_quota"] is not None:
        assert h["threads"] <= max(1, int(h["cgroup_cpu_quota"]))
    assert isinstance(h["model"], str) and h["model"]


@pytest.mark.parametrize("kernel", sorted(bench.PIPE_KERNELS.values()))
def test_pipeline_traffic_is_read_from_the_profile(kernel):
    t = bench.pipeline_traffic(kernel)
    assert t is not None and t > 0
