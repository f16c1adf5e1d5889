"""bench.py's command line on the CPU: the workloads it offers and its refusal of
more than one GPU (the renderer runs on one)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_workloads_listed():
    r = run("--help")
    assert r.returncode == 0
    for w in ("torus", "plain", "heightfield", "icons", "serving"):
        assert w in r.stdout
