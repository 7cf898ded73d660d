"""CPU checks of the measurement contract: bench.py refuses more than one GPU
with a message, and the committed counter summary's per-launch traffic is the
sum its own per-dispatch figures give."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_traffic_summary_is_consistent():
    """bytes_per_launch = (sum of FETCH_SIZE KiB / fetch factor + sum of
    WRITE_SIZE KiB / write factor) over one traversal's dispatches, x 1024,
    / launches (pairs) per traversal."""
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        data = json.load(f)
    for workload, kernels in data.items():
        for name, e in kernels.items():
            n = e["dispatches_per_bfs"]
            fb = sum(e["fetch_kib_per_dispatch"][k] * n[k] for k in n) / e["fetch_factor"]
            wb = sum(e["write_kib_per_dispatch"][k] * n[k] for k in n) / e["write_factor"]
            per = (fb + wb) * 1024 / e["pairs_per_bfs"]
            assert abs(per - e["bytes_per_launch"]) <= 1, (workload, name, per)
            assert os.path.exists(os.path.join(ROOT, e["summary"]))
