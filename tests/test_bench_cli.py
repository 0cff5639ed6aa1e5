"""bench.py contract checks that need no GPU."""
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr or "single" in r.stderr


def test_traffic_profile_is_committed_and_consistent():
    """roofline.traffic comes from the committed rocprofv3 summary."""
    j = json.loads((ROOT / "profiles" / "r3_fp64_n2097152_traffic.json").read_text())
    expect = (j["fetch_kib"] * j["fetch_scale"] + j["write_kib"] * j["write_scale"]) * 1024.0
    assert abs(j["hbm_bytes_per_launch"] - expect) <= 1e-6 * expect
    assert (ROOT / j["source"]).exists()
