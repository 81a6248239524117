"""Caller-defined solvers through the C++ facade's bkma_run / impl_selector and the header-only executors
(parallel-advection_amd/cpp/bkma/plugin.hpp, executors.hpp): the reference's solver plug-in contract
(src/core/bkma_run.hpp:10-13, src/core/AdaptiveWg.hpp:57-140) with two solvers the library does not ship,
checked against closed forms:
  - ShiftSolver: one step circularly shifts every (i0, :, i2) line by s(i0) = i0 % 5 + 1 rows (np.roll);
  - WindowSumSolver(W): row iw < n1-W+1 becomes the sum of rows iw .. iw+W-1, later rows keep their values
    (integer-valued data, so every sum is exact in any order).
The solvers live in tests/cpp/plugin_solvers.hip, built into tests/cpp/libplugin_solvers.so."""
import ctypes
import os

import numpy as np
import pytest

import bkma_amd as B

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "cpp", "libplugin_solvers.so")
ADAPTIVE_WG, ND_RANGE, BASIC_RANGE = 0, 1, 2
_L = None


def _plugin():
    global _L
    if _L is None:
        B.lib()  # the library and torch's HIP runtime first
        if not os.path.exists(LIB):
            pytest.fail(f"{LIB} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB)
        L.plugin_run.restype = ctypes.c_int
        L.plugin_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_ulonglong,
                                 ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_longlong, ctypes.c_int,
                                 ctypes.c_char_p, ctypes.c_int]
        _L = L
    return _L


def _run(solver, impl, t, window=1, global_scratch=0):
    import torch
    torch.cuda.synchronize()
    err = ctypes.create_string_buffer(512)
    n0, n1, n2 = t.shape
    rc = _plugin().plugin_run(solver, impl, int(t.dtype == torch.float64), ctypes.c_void_p(t.data_ptr()), n0, n1,
                              n2, window, global_scratch, err, 512)
    if rc:
        raise RuntimeError(err.value.decode())


def _shift_ref(a, steps):
    out = a.copy()
    for i0 in range(a.shape[0]):
        out[i0] = np.roll(a[i0], steps * (i0 % 5 + 1), axis=0)
    return out


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize("impl,global_scratch", [(ADAPTIVE_WG, 0), (ADAPTIVE_WG, 1), (ND_RANGE, 0),
                                                 (BASIC_RANGE, 0)])
@pytest.mark.parametrize("shape", [(7, 256, 5), (3, 1000, 1)])
def test_shift_solver(cuda, dtype_name, impl, global_scratch, shape):
    import torch
    print(f"shift solver impl {impl} global {global_scratch} {dtype_name} {shape}", flush=True)
    host = np.random.default_rng(5).random(shape).astype(dtype_name)
    t = torch.from_numpy(host).to("cuda")
    for _ in range(3):
        _run(0, impl, t, global_scratch=global_scratch)
    assert np.array_equal(t.cpu().numpy(), _shift_ref(host, 3))


def _window_sum_ref(a, w):
    nw = a.shape[1] - w + 1
    out = a.copy()
    out[:, :nw, :] = sum(a[:, j:j + nw, :] for j in range(w))
    return out


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize("global_scratch", [0, 1])
@pytest.mark.parametrize("shape,w", [((6, 256, 3), 3), ((2, 512, 1), 5)])
def test_window_sum_solver(cuda, dtype_name, global_scratch, shape, w):
    import torch
    print(f"window-sum solver W {w} global {global_scratch} {dtype_name} {shape}", flush=True)
    host = np.random.default_rng(6).integers(0, 8, shape).astype(dtype_name)
    t = torch.from_numpy(host).to("cuda")
    ref = host.copy()
    for _ in range(2):
        _run(1, ADAPTIVE_WG, t, window=w, global_scratch=global_scratch)
        ref = _window_sum_ref(ref, w)
    assert np.array_equal(t.cpu().numpy(), ref)


def test_plugin_errors_reach_the_caller(cuda):
    import torch
    t = torch.zeros((2, 2048, 1), dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="NDRange: n1 = 2048 exceeds the maximum work-group size of 1024"):
        _run(0, ND_RANGE, t)
