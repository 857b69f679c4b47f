"""bench.py contract checks that need no GPU."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_multi_gpu_request_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_help_lists_workloads():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0
    for w in ("16^4", "32^4", "batch", "serve", "precision", "traj32", "heatbath", "hex", "force", "latency",
              "--kernels-only", "--legs", "--no-graphs"):
        assert w in r.stdout


def test_traffic_summary_names_the_roofline_kernels():
    """The committed counter summary carries the HBM bytes per launch that
    the bench line reports as roofline.traffic for each workload's kernel."""
    data = json.load(open(os.path.join(ROOT, "profiles", "kernel_traffic.json")))
    assert data["16^4"]["mix_b_f32"] > 0 and data["32^4"]["mix_b_f16"] > 0
    assert data["32^4"]["force_gauge"] > 0 and data["batch"]["kb_hop"] > 0
