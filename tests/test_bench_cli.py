"""CPU checks of bench.py's contract that need no GPU: a multi-GPU request is refused with a message and a
non-zero status (this library runs one GPU), and the per-launch HBM traffic that the headline line carries
in its roofline objects is recorded, for the kernel the default run launches, in profiles/traffic.json."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_multi_gpu_request_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout.strip() == ""  # no JSON line


def test_headline_traffic_is_recorded():
    import bench
    t = bench.load_traffic("c2_f64", "reg_sweep_f64")
    assert t is not None and t > 0
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        entry = json.load(f)["c2_f64"]
    assert entry["bytes_per_launch"] == t
    # the traffic is the scaled sum of the two counters (KiB per dispatch)
    scaled = (entry["fetch_kib_per_dispatch"] * entry["fetch_scale"]
              + entry["write_kib_per_dispatch"] * entry["write_scale"]) * 1024
    assert abs(scaled - t) / t < 1e-3
    assert bench.load_traffic("c2_f64", "some_other_kernel") is None
