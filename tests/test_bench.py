"""bench.py's host-side contract: the algorithmic byte counts behind its roofline figures and the
one-GPU limit (no GPU needed)."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


@pytest.mark.parametrize("n, gb", [(128, 0.0999), (192, 0.3368), (256, 0.7980), (384, 2.692),
                                   (512, 6.3795), (1024, 51.02)])
def test_step_bytes_match_the_metric_definition(n, gb):
    # 60 N_p + 32 N^3 + 16 N^2 (N/2 + 1) bytes per FP32 step with N_p = (N/2)^3
    b = bench.step_bytes(n, (n // 2) ** 3, 32)
    assert abs(b / 1e9 - gb) / gb < 5e-4
    assert bench.step_bytes(n, (n // 2) ** 3, 64) == 2 * b


def test_stage_bytes_of_the_workload():
    n, n_p = 512, 256 ** 3
    sb = bench.stage_bytes(n, n_p, 32)
    half = n * n * (n // 2 + 1) * 8
    assert sb["c2c_y_fwd"] == sb["z_green"] == sb["c2c_y_inv"] == 2 * half
    assert sb["r2c_x"] == sb["c2r_x"] == 4 * n ** 3 + half
    assert sb["gather_kick_drift"] == 52 * n_p + 4 * n ** 3      # 1.41 GB
    assert sb["deposit"] == 16 * n_p + 4 * n ** 3                # 805 MB
    assert abs(sb["gather_kick_drift"] / 1e9 - 1.409) < 1e-3


def test_loop_steps_replay_the_reference_loop():
    assert bench.loop_steps(0.02, 1.0, 1000) == 1000
    assert bench.loop_steps(0.02, 1.0, 100) == 99


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU only" in (r.stdout + r.stderr)
