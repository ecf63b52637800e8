"""bench.py's command-line contract on the CPU: it measures one GPU, so any
other --gpus value exits with status 2 and a message before touching the
library or a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "8"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stdout


def test_bench_help_lists_the_schedule_options():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--panel-width", "--gemm-tile"):
        assert opt in r.stdout
