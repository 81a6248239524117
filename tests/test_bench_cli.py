"""bench.py's command-line contract, checked without a GPU: more than one GPU
and an empty timed region are refused before anything is measured."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=300, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_zero_timed_steps_is_refused():
    r = _bench("--steps", "0")
    assert r.returncode == 2
    assert "--steps must be at least 1" in r.stderr
