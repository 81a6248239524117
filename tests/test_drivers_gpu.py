"""The drop-in drivers parallel-advection_amd/bin/advection and bin/conv1d (cpp/apps/), the counterparts of the
reference's src/advection.cpp and src/conv1d.cpp, run on reduced .ini files: their RESULTS_VALIDATION and
"Normalized Array" outputs against the oracle, their PERF_DIAGS lines, and impl_selector's error for a name the
reference rejects.  The drivers print with std::cout's default six significant digits, which bounds the
comparisons."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "parallel-advection_amd", "bin")

ADV_INI = """[problem]
n0 = {n0}
n1 = {n1}
n2 = {n2}
maxIter = {steps}
dt  = 0.001
minRealX  = 0
maxRealX  = 1
minRealVx = -1
maxRealVx = 1

[impl]
kernelImpl  = {impl}
inplace = true

[optimization]
gpu     = true
pref_wg_size = 512
seq_size0 = 1
seq_size2 = 1

[io]
outputSolution = false
"""

CONV_INI = """[problem]
length = {length}
channel_in = {c}
k = {k}
total_batch_size = {total}
batch_size_n2 = {n2}

[impl]
kernelImpl  = AdaptiveWg
inplace = true

[optimization]
gpu     = true
pref_wg_size = 512
seq_size0 = 1
seq_size2 = 1
"""


def _run(binary, text, tmp_path):
    ini = tmp_path / "run.ini"
    ini.write_text(text)
    exe = os.path.join(BIN, binary)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is missing: build it with __graft_entry__.build()")
    print(f"driver {binary} on {ini}", flush=True)
    return subprocess.run([exe, str(ini)], capture_output=True, text=True, timeout=120)


def _perf(out):
    keys = {}
    for k, unit in (("elapsed_time", "s"), ("upd_cells_per_sec", "Gcell/sec"), ("estimated_throughput", "GB/s")):
        m = re.search(rf"^{k}: (\S+) {re.escape(unit)}$", out, flags=re.M)
        assert m, k
        keys[k] = float(m.group(1))
    return keys


@pytest.mark.parametrize("impl", ["AdaptiveWg", "NDRange", "Tuned"])
def test_advection_driver_validation_matches_oracle(cuda, tmp_path, impl):
    n0, n1, n2, steps = 64, 1024, 1, 20
    r = _run("advection", ADV_INI.format(n0=n0, n1=n1, n2=n2, steps=steps, impl=impl), tmp_path)
    assert r.returncode == 0, r.stderr
    assert "PERF_DIAGS:" in r.stdout and "RESULTS_VALIDATION:" in r.stdout
    m = re.search(r"Highest L1 error found: (\S+) \(lowest is (\S+)\)", r.stdout)
    assert m, r.stdout
    hi, lo = float(m.group(1)), float(m.group(2))
    # dt arrives float-rounded through the .ini reader (ConfigMap::getFloat)
    p = O.make_params(n0, n1, n2, dt=float(np.float32(0.001)), minRealVx=-1, maxRealVx=1, maxIter=steps)
    f = O.fill_adv_sin4(p)
    O.advect_f64(f, p, steps)
    ohi, olo = O.l1_error_adv(f, p)
    print(f"{impl}: driver L1 {hi:.6g} / {lo:.6g}, oracle {ohi:.6g} / {olo:.6g}", flush=True)
    assert abs(hi - ohi) <= 5e-6 * ohi and abs(lo - olo) <= 5e-6 * olo
    assert hi < 1e-6  # the reference tests' analytic criterion
    perf = _perf(r.stdout)
    cells = n0 * n1 * n2 * steps
    assert abs(perf["upd_cells_per_sec"] - cells / perf["elapsed_time"] / 1e9) <= 1e-4 * perf["upd_cells_per_sec"]
    assert abs(perf["estimated_throughput"] - 16 * perf["upd_cells_per_sec"]) <= 1e-4 * perf["estimated_throughput"]


def test_advection_driver_rejects_names_the_reference_rejects(cuda, tmp_path):
    # a shape the reference's create_optim_params accepts (it runs before impl_selector in the driver)
    r = _run("advection", ADV_INI.format(n0=16, n1=1024, n2=1, steps=1, impl="BasicRange"), tmp_path)
    assert r.returncode == 1
    assert ("BasicRange is not a valid implementation name.\nShould be: {BasicRange, NDRange, AdaptiveWg}"
            in r.stderr)


def test_conv1d_driver_matches_oracle(cuda, tmp_path):
    length, c, k, total, n2 = 64, 3, 3, 512, 8
    n0, n1, n_write = total // n2, length * c, length - k + 1
    r = _run("conv1d", CONV_INI.format(length=length, c=c, k=k, total=total, n2=n2), tmp_path)
    assert r.returncode == 0, r.stderr
    before = float(re.search(r"Normalized Array before: (\S+)", r.stdout).group(1))
    after = float(re.search(r"Normalized Array after: (\S+)", r.stdout).group(1))
    assert "All values data[:,i1,:] are equal." in r.stdout
    data = np.full((n0, n1, n2), 7.3)
    O.conv_literal_f64(data, np.full((k, c, c), 1.5), np.ones(c), k, c, length)
    expect = data[:, :n_write, :].mean()
    print(f"conv1d driver: before {before}, after {after}, oracle {expect}", flush=True)
    assert abs(before - 7.3) <= 5e-6 * 7.3
    assert abs(after - expect) <= 5e-6 * expect
    assert abs(expect - (1 + c * k * 1.5 * 7.3)) < 1e-9  # the reference's constant-input known answer, 99.55
    _perf(r.stdout)
