"""bench.py's command-line contract that needs no GPU: more than one GPU is
refused with a message (the simulation step of one network is not split
across GPUs in this build), and the workload list names the headline one."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""  # no JSON line


def test_workloads():
    sys.path.insert(0, ROOT)
    import bench
    assert {"large_1m", "plastic_100k", "balanced_10k", "brunel_event"} <= set(bench.WORKLOADS)
    assert bench.WORKLOAD_N["large_1m"] == 1_000_000
