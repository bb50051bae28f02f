"""bench.py on the CPU: it refuses more than one GPU before touching one, and
the `traffic` figure of each benchmark workload comes from the committed
rocprofv3 summary, within a few percent of the workload's algorithmic bytes."""
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_workload_traffic_comes_from_the_committed_summary():
    table = json.loads((ROOT / "profiles" / "hbm_traffic.json").read_text())["workloads"]
    sizes = {"F64": (768, 8), "F32": (512, 4), "F32DEV": (512, 4), "F16DEV": (512, 2), "BF16DEV": (512, 2),
             "I16DEV": (512, 2)}
    for form, (n, el) in sizes.items():
        tau = "0.5015" if form == "F64" else "0.505"
        wl = f"D3Q27 cumulant A-A {form}, {n}^3 periodic box, tau={tau}"
        alg = 2 * 27 * el * n ** 3
        assert table[wl]["algorithmic_bytes_per_launch"] == alg
        assert alg <= bench.measured_traffic(wl) <= 1.02 * alg
    assert bench.measured_traffic("no such workload") is None


def test_placement_policy_is_a_command_line_choice():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--help"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    assert "--placement" in r.stdout and "--prec-warmup" in r.stdout
