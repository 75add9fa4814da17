"""bench.py's host logic on the CPU: the one-GPU contract and the committed
profile figures it copies into its JSON line."""
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


def test_time_to_target_figures_are_read_from_the_profile():
    t = bench.committed_time_to_2pct()
    assert t is not None
    assert t["mc_plus_denoiser_s"] < t["plain_gpu_mc_s"] < t["cpu_oracle_s"]
    assert t["speedup_mc_plus_denoiser"] == pytest.approx(t["plain_gpu_mc_s"] / t["mc_plus_denoiser_s"])
    assert t["speedup_mc_plus_denoiser"] >= 5.0


@pytest.mark.parametrize("kernel", sorted(bench.PIPE_KERNELS.values()))
def test_pipeline_traffic_is_read_from_the_profile(kernel):
    t = bench.pipeline_traffic(kernel)
    assert t is not None and t > 0
