"""bench.py measures one GPU: asked for more it exits with status 2, says
why on stderr and prints no result line (it returns before touching a
device, so this runs without one)."""
import os
import subprocess
import sys

ROOT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..")


def test_bench_refuses_more_than_one_gpu():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--gpus", "2"], cwd=ROOT, capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and "--gpus 2" in p.stderr
    assert p.stdout.strip() == ""
