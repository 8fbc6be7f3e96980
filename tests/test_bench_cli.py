"""bench.py's command line without a GPU: it runs on one GPU only and says so
for any other count, before it touches a device."""
import json
import os
import subprocess
import sys
from pathlib import Path

REPO = Path(__file__).resolve().parent.parent


def run_bench(*args, env=None):
    return subprocess.run([sys.executable, str(REPO / "bench.py"), *args], cwd=REPO, capture_output=True,
                          text=True, timeout=120, env=env)


def test_more_than_one_gpu_is_refused_with_a_message():
    p = run_bench("--gpus", "2")
    assert p.returncode == 2
    assert "one GPU only" in p.stderr
    assert p.stdout == ""


def test_world_size_above_one_is_refused():
    env = dict(os.environ, WORLD_SIZE="4")
    p = run_bench(env=env)
    assert p.returncode == 2
    assert "WORLD_SIZE=4" in p.stderr


def test_traffic_file_of_the_default_workload_is_committed():
    """The roofline's traffic figure comes from the committed rocprofv3
    counters of the default workload."""
    t = json.loads((REPO / "profiles" / "traffic_ellipsoids512.json").read_text())
    ka = t["k_pass_a"]
    assert abs(ka["hbm_bytes_per_launch"] - ka["hbm_read_bytes_per_launch"] - ka["hbm_write_bytes_per_launch"]) <= 1
    assert ka["hbm_bytes_per_launch"] > ka["algorithmic_bytes_per_launch"] > 0
    assert ka["valu_insts_per_launch"] > 0
    # keyed to the kernel sources it was measured on (bench.py compares the key)
    assert len(t["kernel_sources_sha256"]) == 64 and t["kernel_sources"]
