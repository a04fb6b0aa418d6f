"""bench.py's command-line contract on the host: one GPU only (several GPUs
exit with a message), and the host description it reports."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_host_info():
    sys.path.insert(0, ROOT)
    import bench
    h = bench.host_info()
    assert h["nproc"] == os.cpu_count()
    assert set(h) == {"nproc", "cpu_model", "cpu_mhz"}
