"""bench.py's command-line contract on the host: more than one GPU is refused with a message
and exit status 2, before any GPU work; the copy-calibration and eigen options parse."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_bench_help_lists_options():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--gpus", "--eigen", "--eigen-budget", "--copy-calibration", "--no-cpu"):
        assert opt in r.stdout
