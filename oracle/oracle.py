"""ctypes front end of the CPU oracle (liboracle.so built from bkma_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker -- never by the product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


class AdvParams(ctypes.Structure):
    """Same field layout as bkma_adv_params (mirrors the reference ADVParams,
    src/config/AdvectionParams.hpp:10-57)."""

    _fields_ = [
        ("n0", ctypes.c_uint64), ("n1", ctypes.c_uint64), ("n2", ctypes.c_uint64),
        ("maxIter", ctypes.c_uint64), ("pref_wg_size", ctypes.c_uint64),
        ("seq_size0", ctypes.c_uint64), ("seq_size2", ctypes.c_uint64),
        ("dt", ctypes.c_double), ("dx", ctypes.c_double), ("dvx", ctypes.c_double),
        ("inv_dx", ctypes.c_double), ("minRealX", ctypes.c_double),
        ("maxRealX", ctypes.c_double), ("realWidthX", ctypes.c_double),
        ("minRealVx", ctypes.c_double), ("maxRealVx", ctypes.c_double),
        ("percent_loc", ctypes.c_float), ("gpu", ctypes.c_int32),
        ("outputSolution", ctypes.c_int32), ("_pad", ctypes.c_int32),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        dp, fp = P(ctypes.c_double), P(ctypes.c_float)
        u64 = ctypes.c_uint64
        L.orc_uniform.restype = ctypes.c_double
        L.orc_uniform.argtypes = [u64, u64]
        L.orc_fill_uniform_f64.argtypes = [dp, u64, u64]
        L.orc_fill_uniform_f32.argtypes = [fp, u64, u64]
        L.orc_update_deltas.argtypes = [P(AdvParams)]
        L.orc_lag_basis.argtypes = [ctypes.c_double, dp]
        L.orc_displ.restype = ctypes.c_double
        L.orc_displ.argtypes = [P(AdvParams), ctypes.c_int, ctypes.c_int]
        L.orc_adv_foot.restype = ctypes.c_int
        L.orc_adv_foot.argtypes = [P(AdvParams), ctypes.c_int, ctypes.c_int, dp]
        L.orc_advect_f64.argtypes = [dp, P(AdvParams), u64]
        L.orc_advect_f32_mixed.argtypes = [fp, P(AdvParams), u64]
        L.orc_advect_f32_plain.argtypes = [fp, P(AdvParams), u64]
        L.orc_advect_f32_blocked.argtypes = [fp, P(AdvParams), u64, u64]
        L.orc_fill_adv_sin4.argtypes = [dp, P(AdvParams)]
        L.orc_l1_error_adv.restype = ctypes.c_double
        L.orc_l1_error_adv.argtypes = [dp, P(AdvParams), dp]
        L.orc_conv_literal_f64.argtypes = [dp, u64, u64, u64, dp, dp, u64, u64, u64]
        L.orc_conv_torch_f64.argtypes = [dp, u64, u64, u64, dp, dp, u64]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_abi_sizeof_params.restype = ctypes.c_int
        assert L.orc_abi_sizeof_params() == ctypes.sizeof(AdvParams)
        _lib = L
    return _lib


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def make_params(n0, n1, n2, dt=1e-4, minRealX=0.0, maxRealX=1.0, minRealVx=0.0, maxRealVx=1.0,
                maxIter=100, pref_wg_size=128, seq_size0=1, seq_size2=1) -> AdvParams:
    """ADVParams with update_deltas applied (AdvectionParams.cpp:96-103)."""
    p = AdvParams(n0=n0, n1=n1, n2=n2, maxIter=maxIter, pref_wg_size=pref_wg_size,
                  seq_size0=seq_size0, seq_size2=seq_size2, dt=dt, minRealX=minRealX,
                  maxRealX=maxRealX, minRealVx=minRealVx, maxRealVx=maxRealVx,
                  percent_loc=1.0, gpu=0, outputSolution=0)
    lib().orc_update_deltas(ctypes.byref(p))
    return p


def uniform(n: int, seed: int, dtype=np.float64) -> np.ndarray:
    out = np.empty(n, dtype=dtype)
    if dtype == np.float64:
        lib().orc_fill_uniform_f64(_dptr(out), n, seed)
    else:
        lib().orc_fill_uniform_f32(_fptr(out), n, seed)
    return out


def lag_basis(px: float) -> np.ndarray:
    c = np.zeros(6)
    lib().orc_lag_basis(px, _dptr(c))
    return c


def adv_foot(p: AdvParams, i0: int, i1: int):
    d = ctypes.c_double()
    ln = lib().orc_adv_foot(ctypes.byref(p), i0, i1, ctypes.byref(d))
    return ln, d.value


def advect_f64(data: np.ndarray, p: AdvParams, nsteps: int) -> None:
    lib().orc_advect_f64(_dptr(data), ctypes.byref(p), nsteps)


def advect_f32_mixed(data: np.ndarray, p: AdvParams, nsteps: int) -> None:
    lib().orc_advect_f32_mixed(_fptr(data), ctypes.byref(p), nsteps)


def advect_f32_blocked(data: np.ndarray, p: AdvParams, nsteps: int, steps_per_pass: int) -> None:
    lib().orc_advect_f32_blocked(_fptr(data), ctypes.byref(p), nsteps, steps_per_pass)


def advect_f32_plain(data: np.ndarray, p: AdvParams, nsteps: int) -> None:
    lib().orc_advect_f32_plain(_fptr(data), ctypes.byref(p), nsteps)


def fill_adv_sin4(p: AdvParams) -> np.ndarray:
    out = np.empty((p.n0, p.n1, p.n2))
    lib().orc_fill_adv_sin4(_dptr(out), ctypes.byref(p))
    return out


def l1_error_adv(data: np.ndarray, p: AdvParams):
    lo = ctypes.c_double()
    hi = lib().orc_l1_error_adv(_dptr(data), ctypes.byref(p), ctypes.byref(lo))
    return hi, lo.value


def conv_literal_f64(data: np.ndarray, w_kco: np.ndarray, b: np.ndarray, K: int, C: int, L: int) -> None:
    n0, n1, n2 = data.shape
    lib().orc_conv_literal_f64(_dptr(data), n0, n1, n2, _dptr(w_kco), _dptr(b), K, C, L)


def conv_torch_f64(x: np.ndarray, w_ock: np.ndarray, b: np.ndarray) -> None:
    B, C, L = x.shape
    K = w_ock.shape[2]
    lib().orc_conv_torch_f64(_dptr(x), B, C, L, _dptr(w_ock), _dptr(b), K)


def num_threads() -> int:
    return lib().orc_num_threads()
