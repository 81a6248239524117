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


def _bench_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_unknown_config_is_refused():
    r = _bench("--configs", "c1,c9")
    assert r.returncode == 2
    assert "unknown --configs" in r.stderr


def test_headline_roofline_counts_the_bytes_a_launch_moves():
    # The headline kernel fuses several steps per launch; its roofline divides the bytes the launch must move
    # (each FP32 cell read and written once) by the launch time, so the fraction cannot exceed 1 below the
    # peak bandwidth.  Figures of the headline workload: 2048*1024*512 cells, 2.433 ms per 5-step launch.
    b = _bench_module()
    cells = 2048 * 1024 * 512
    r = b.roofline("hbm", cells * 2 * 4, 2.433, b.HBM_PEAK_GBS, "GB/s")
    assert abs(r["achieved"] - cells * 8 / 2.433e-3 / 1e9) < 0.01
    assert 0.43 < r["frac"] < 0.45
    # a launch as fast as the peak allows gives exactly 1
    assert b.roofline("hbm", 8e9, 1.0, 8000.0, "GB/s")["frac"] == 1.0
    # MFMA rooflines count flop in TFLOP/s
    c2 = b.roofline("mfma", 6.865e11, 0.4576, b.BF16_DENSE_PEAK_TFLOPS, "TFLOP/s")
    assert abs(c2["frac"] - 0.6) < 0.001
