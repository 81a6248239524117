"""Python binding of libbkma_amd.so (the gfx950 BKMA library) over its C ABI.

The product is the C ABI declared in include/bkma_amd.h and the C++ facade in
parallel-advection_amd/cpp; this module is the thin ctypes layer the tests and
bench.py use.  Arrays are torch tensors already resident on the GPU; only
their device pointers cross the boundary.  There is no CPU fallback: if the
shared library is missing, importing the binding raises.
"""
from __future__ import annotations

import ctypes
import os

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_PKG_DIR), "lib", "libbkma_amd.so")

# ---------------------------------------------------------------------------
# Enums (include/bkma_amd.h)
# ---------------------------------------------------------------------------
BKMA_OK = 0
BKMA_ERR_INVALID_ARGUMENT = 1
BKMA_ERR_RUNTIME = 2
BKMA_ERR_DEVICE = 3
BKMA_ERR_UNSUPPORTED = 4
BKMA_ERR_OUT_OF_MEMORY = 5

BASIC_RANGE, ND_RANGE, ADAPTIVE_WG = 0, 1, 2
MEM_LOCAL, MEM_GLOBAL = 0, 1
F64, F32, BF16 = 0, 1, 2
CONV_F64, CONV_F32, CONV_BF16, CONV_BF16X3, CONV_FP8, CONV_BF16X6 = 0, 1, 2, 3, 4, 5


class BkmaError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(msg)
        self.status = status


class BkmaInvalidArgument(BkmaError, ValueError):
    """std::invalid_argument in the reference."""


class BkmaUnsupported(BkmaError):
    pass


class BkmaDeviceError(BkmaError):
    pass


_EXC = {BKMA_ERR_INVALID_ARGUMENT: BkmaInvalidArgument, BKMA_ERR_UNSUPPORTED: BkmaUnsupported,
        BKMA_ERR_DEVICE: BkmaDeviceError}


# ---------------------------------------------------------------------------
# Structs
# ---------------------------------------------------------------------------
class AdvParams(ctypes.Structure):
    """bkma_adv_params == the reference's ADVParams (AdvectionParams.hpp:10-57)."""

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

    def update_deltas(self) -> "AdvParams":
        lib().bkma_adv_update_deltas(ctypes.byref(self))
        return self


class BatchConfig1D(ctypes.Structure):
    _fields_ = [("n_batch", ctypes.c_uint64), ("batch_size", ctypes.c_uint64),
                ("last_batch_size", ctypes.c_uint64)]


class WgDispatch(ctypes.Structure):
    _fields_ = [("s0", ctypes.c_uint64), ("s2", ctypes.c_uint64), ("g0", ctypes.c_uint64),
                ("g2", ctypes.c_uint64)]


class OptimParams(ctypes.Structure):
    """bkma_optim_params == BkmaOptimParams (bkma_tools.hpp:126-134)."""

    _fields_ = [("dispatch_d0", BatchConfig1D), ("dispatch_d2", BatchConfig1D),
                ("w0", ctypes.c_uint64), ("w1", ctypes.c_uint64), ("w2", ctypes.c_uint64),
                ("wg_dispatch", WgDispatch), ("mem_space", ctypes.c_int32), ("_pad", ctypes.c_int32)]


class TunedConfig(ctypes.Structure):
    """bkma_tuned_config."""

    _fields_ = [("strategy", ctypes.c_int32), ("vec", ctypes.c_int32), ("smax", ctypes.c_int32),
                ("prefetch", ctypes.c_int32), ("waves_per_wg", ctypes.c_int32),
                ("nontemporal", ctypes.c_int32), ("persistent", ctypes.c_int32), ("seg_rows", ctypes.c_int32),
                ("wide_window", ctypes.c_int32), ("steps_per_pass", ctypes.c_int32),
                ("chunk_segs", ctypes.c_int32), ("dpp_broadcast", ctypes.c_int32)]


STRATEGY_COLUMN_STREAM = 1
STRATEGY_LINE_LDS = 2
STRATEGY_COLUMN_PAIR = 3
STRATEGY_PLAIN = 4
STRATEGY_HYBRID = 5
STRATEGY_LINE_SHFL = 6


class AdvProblem(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("params", AdvParams)]


class Span3d(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("e0", ctypes.c_uint64), ("e1", ctypes.c_uint64),
                ("e2", ctypes.c_uint64)]


# ---------------------------------------------------------------------------
# Library loading
# ---------------------------------------------------------------------------
_lib = None

EXPORTS = [
    "bkma_abi_version", "bkma_last_error", "bkma_query_lds_bytes", "bkma_adv_update_deltas",
    "bkma_create_optim_params", "bkma_init_1d_blocking", "bkma_scratch_extents", "bkma_adv_run",
    "bkma_conv_run", "bkma_fill_uniform", "bkma_fill_normal", "bkma_fill_constant", "bkma_fill_adv_sin4",
    "bkma_validate_adv_l1", "bkma_conv_sum_normalize", "bkma_conv_validate", "bkma_error_norms",
    "bkma_adv_plan_create", "bkma_adv_plan_create_with", "bkma_adv_plan_describe", "bkma_adv_plan_run",
    "bkma_adv_plan_destroy", "bkma_adv_plan_variants", "bkma_adv_batch_create", "bkma_adv_batch_create_with",
    "bkma_adv_batch_run",
    "bkma_adv_batch_run_loop", "bkma_adv_batch_describe", "bkma_adv_batch_steps_per_pass", "bkma_adv_batch_destroy",
    "bkma_conv1d_plan_create", "bkma_conv1d_plan_run", "bkma_conv1d_plan_describe", "bkma_conv1d_plan_destroy",
    "bkma_conv1d_plan_set_input_scale", "bkma_adv_run_noalloc", "bkma_static_scratch_bytes",
]


def lib():
    """Load libbkma_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          f"(no CPU fallback exists for the BKMA kernels)")
    # torch ships its own HIP runtime under the same soname (libamdhip64.so.7).
    # Load it first so that libbkma_amd.so binds to that copy: two HIP
    # runtimes in one process do not see the same devices.
    import torch  # noqa: F401
    L = ctypes.CDLL(LIB_PATH)
    P, u64, i32, vp, dbl = ctypes.POINTER, ctypes.c_uint64, ctypes.c_int32, ctypes.c_void_p, ctypes.c_double
    st = ctypes.c_int
    sig = {
        "bkma_abi_version": (ctypes.c_int, []),
        "bkma_last_error": (ctypes.c_char_p, []),
        "bkma_query_lds_bytes": (st, [ctypes.c_int, P(u64)]),
        "bkma_adv_update_deltas": (None, [P(AdvParams)]),
        "bkma_create_optim_params": (st, [u64, u64, u64, u64, u64, u64, u64, u64, P(OptimParams)]),
        "bkma_init_1d_blocking": (BatchConfig1D, [u64, u64]),
        "bkma_scratch_extents": (st, [ctypes.c_int, P(OptimParams), u64, u64, u64, u64, P(u64), P(u64), P(u64)]),
        "bkma_adv_run": (st, [ctypes.c_int, ctypes.c_int, vp, P(AdvParams), P(OptimParams), Span3d, vp]),
        "bkma_adv_run_noalloc": (st, [ctypes.c_int, ctypes.c_int, vp, P(AdvParams), P(OptimParams), vp]),
        "bkma_static_scratch_bytes": (u64, []),
        "bkma_conv_run": (st, [ctypes.c_int, ctypes.c_int, vp, u64, u64, u64, vp, vp, u64, u64, u64,
                               P(OptimParams), Span3d, vp]),
        "bkma_fill_uniform": (st, [ctypes.c_int, vp, u64, u64, vp]),
        "bkma_fill_normal": (st, [ctypes.c_int, vp, u64, u64, dbl, vp]),
        "bkma_fill_constant": (st, [ctypes.c_int, vp, u64, dbl, vp]),
        "bkma_fill_adv_sin4": (st, [ctypes.c_int, vp, P(AdvParams), vp]),
        "bkma_validate_adv_l1": (st, [ctypes.c_int, vp, P(AdvParams), P(dbl), P(dbl), vp]),
        "bkma_conv_sum_normalize": (st, [ctypes.c_int, vp, u64, u64, u64, u64, P(dbl), vp]),
        "bkma_conv_validate": (st, [ctypes.c_int, vp, u64, u64, u64, u64, P(i32), vp]),
        "bkma_error_norms": (st, [ctypes.c_int, vp, ctypes.c_int, vp, u64, P(dbl), vp]),
        "bkma_adv_plan_create": (st, [P(AdvParams), ctypes.c_int, i32, vp, P(vp)]),
        "bkma_adv_plan_create_with": (st, [P(AdvParams), ctypes.c_int, P(TunedConfig), P(vp)]),
        "bkma_adv_plan_describe": (st, [vp, P(TunedConfig), ctypes.c_char_p, u64]),
        "bkma_adv_plan_run": (st, [vp, vp, u64, vp]),
        "bkma_adv_plan_destroy": (None, [vp]),
        "bkma_adv_plan_variants": (i32, [ctypes.c_int, P(TunedConfig), i32]),
        "bkma_adv_batch_create": (st, [P(AdvProblem), i32, ctypes.c_int, i32, vp, P(vp)]),
        "bkma_adv_batch_create_with": (st, [P(AdvProblem), i32, ctypes.c_int, u64, i32, i32, vp, P(vp)]),
        "bkma_adv_batch_run": (st, [vp, u64, i32, vp]),
        "bkma_adv_batch_run_loop": (st, [vp, u64, i32, vp]),
        "bkma_adv_batch_describe": (st, [vp, P(u64), P(u64), P(u64), P(u64), P(ctypes.c_uint32)]),
        "bkma_adv_batch_steps_per_pass": (i32, [vp]),
        "bkma_adv_batch_destroy": (None, [vp]),
        "bkma_conv1d_plan_create": (st, [ctypes.c_int, ctypes.c_int, u64, u64, u64, u64, ctypes.c_int, vp, vp, vp,
                                         P(vp)]),
        "bkma_conv1d_plan_run": (st, [vp, vp, vp]),
        "bkma_conv1d_plan_describe": (st, [vp, ctypes.c_char_p, u64]),
        "bkma_conv1d_plan_destroy": (None, [vp]),
        "bkma_conv1d_plan_set_input_scale": (st, [vp, ctypes.c_float]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(status: int) -> None:
    if status != BKMA_OK:
        msg = lib().bkma_last_error().decode(errors="replace")
        raise _EXC.get(status, BkmaError)(status, msg)


# ---------------------------------------------------------------------------
# Helpers over torch tensors
# ---------------------------------------------------------------------------
def _torch():
    import torch
    return torch


def dtype_code(t) -> int:
    torch = _torch()
    return {torch.float64: F64, torch.float32: F32, torch.bfloat16: BF16}[t.dtype]


def _ptr(t):
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous(), "BKMA arrays are contiguous device tensors"
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def make_params(n0, n1, n2, dt=1e-4, minRealX=0.0, maxRealX=1.0, minRealVx=0.0, maxRealVx=1.0,
                maxIter=100, pref_wg_size=128, seq_size0=1, seq_size2=1) -> AdvParams:
    p = AdvParams(n0=n0, n1=n1, n2=n2, maxIter=maxIter, pref_wg_size=pref_wg_size, seq_size0=seq_size0,
                  seq_size2=seq_size2, dt=dt, minRealX=minRealX, maxRealX=maxRealX, minRealVx=minRealVx,
                  maxRealVx=maxRealVx, percent_loc=1.0, gpu=1, outputSolution=0)
    return p.update_deltas()


def query_lds_bytes(device: int = 0) -> int:
    v = ctypes.c_uint64()
    check(lib().bkma_query_lds_bytes(device, ctypes.byref(v)))
    return v.value


def create_optim_params(n0, n1, n2, pref_wg_size=128, seq_size0=1, seq_size2=1, elem_size=8,
                        local_mem_bytes=0) -> OptimParams:
    o = OptimParams()
    check(lib().bkma_create_optim_params(n0, n1, n2, pref_wg_size, seq_size0, seq_size2, elem_size,
                                         local_mem_bytes, ctypes.byref(o)))
    return o


def scratch_extents(impl, optim, n0, n1, n2, window):
    e = [ctypes.c_uint64() for _ in range(3)]
    check(lib().bkma_scratch_extents(impl, ctypes.byref(optim) if optim is not None else None, n0, n1, n2,
                                     window, *[ctypes.byref(x) for x in e]))
    return tuple(x.value for x in e)


def _span(scratch, extents=None):
    if scratch is None:
        return Span3d(None, 0, 0, 0)
    e = extents if extents is not None else tuple(scratch.shape)
    return Span3d(ctypes.c_void_p(scratch.data_ptr()), *e)


def adv_run(impl, data, params, optim=None, scratch=None, scratch_extents=None, stream=None) -> None:
    """bkma_run<AdvectionSolver, impl>: one in-place step on `data`."""
    check(lib().bkma_adv_run(impl, dtype_code(data), _ptr(data), ctypes.byref(params),
                             ctypes.byref(optim) if optim is not None else None,
                             _span(scratch, scratch_extents), _stream(stream)))


def adv_run_noalloc(impl, data, params, optim=None, stream=None) -> None:
    """bkma_adv_run_noalloc: one in-place step; any global scratch comes from the library's static array."""
    check(lib().bkma_adv_run_noalloc(impl, dtype_code(data), _ptr(data), ctypes.byref(params),
                                     ctypes.byref(optim) if optim is not None else None, _stream(stream)))


def conv_run(impl, data, weight, bias, K, C, L, optim=None, scratch=None, scratch_extents=None,
             stream=None) -> None:
    """bkma_run<ConvSolver, impl> (reference ConvSolver semantics)."""
    n0, n1, n2 = data.shape
    check(lib().bkma_conv_run(impl, dtype_code(data), _ptr(data), n0, n1, n2, _ptr(weight), _ptr(bias), K, C, L,
                              ctypes.byref(optim) if optim is not None else None, _span(scratch, scratch_extents),
                              _stream(stream)))


def fill_uniform(t, seed: int, stream=None) -> None:
    check(lib().bkma_fill_uniform(dtype_code(t), _ptr(t), t.numel(), seed, _stream(stream)))


def fill_normal(t, seed: int, stddev: float = 1.0, stream=None) -> None:
    check(lib().bkma_fill_normal(dtype_code(t), _ptr(t), t.numel(), seed, stddev, _stream(stream)))


def fill_constant(t, value: float, stream=None) -> None:
    check(lib().bkma_fill_constant(dtype_code(t), _ptr(t), t.numel(), value, _stream(stream)))


def fill_adv_sin4(t, params, stream=None) -> None:
    check(lib().bkma_fill_adv_sin4(dtype_code(t), _ptr(t), ctypes.byref(params), _stream(stream)))


def validate_adv_l1(t, params, stream=None):
    hi, lo = ctypes.c_double(), ctypes.c_double()
    check(lib().bkma_validate_adv_l1(dtype_code(t), _ptr(t), ctypes.byref(params), ctypes.byref(hi),
                                     ctypes.byref(lo), _stream(stream)))
    return hi.value, lo.value


def error_norms(a, ref, stream=None):
    """(max|a-ref|, max|ref|, sum (a-ref)^2, sum ref^2) in FP64."""
    out = (ctypes.c_double * 4)()
    assert a.numel() == ref.numel()
    check(lib().bkma_error_norms(dtype_code(a), _ptr(a), dtype_code(ref), _ptr(ref), a.numel(), out,
                                 _stream(stream)))
    return tuple(out)


def rel_errors(a, ref, stream=None):
    """rel_max = max|a-ref|/max|ref|, rel_L2 = ||a-ref||_2/||ref||_2."""
    mx, mr, s2, r2 = error_norms(a, ref, stream)
    return mx / mr if mr else mx, (s2 ** 0.5) / (r2 ** 0.5) if r2 else s2 ** 0.5


def conv_sum_normalize(t, nw, stream=None) -> float:
    n0, n1, n2 = t.shape
    m = ctypes.c_double()
    check(lib().bkma_conv_sum_normalize(dtype_code(t), _ptr(t), n0, n1, n2, nw, ctypes.byref(m), _stream(stream)))
    return m.value


def conv_validate(t, nw, stream=None) -> bool:
    n0, n1, n2 = t.shape
    f = ctypes.c_int32()
    check(lib().bkma_conv_validate(dtype_code(t), _ptr(t), n0, n1, n2, nw, ctypes.byref(f), _stream(stream)))
    return bool(f.value)


def plan_variants(dtype):
    """The compiled tuned-advection configurations for `dtype`."""
    n = lib().bkma_adv_plan_variants(dtype, None, 0)
    arr = (TunedConfig * n)()
    lib().bkma_adv_plan_variants(dtype, arr, n)
    return list(arr)


class AdvPlan:
    """A tuned in-place advection plan (bkma_adv_plan_*)."""

    def __init__(self, params, dtype, autotune: bool = False, config: TunedConfig | None = None, stream=None):
        self._h = ctypes.c_void_p()
        self.params = params
        if config is None:
            check(lib().bkma_adv_plan_create(ctypes.byref(params), dtype, int(autotune), _stream(stream),
                                             ctypes.byref(self._h)))
        else:
            check(lib().bkma_adv_plan_create_with(ctypes.byref(params), dtype, ctypes.byref(config),
                                                  ctypes.byref(self._h)))

    def run(self, data, nsteps: int = 1, stream=None) -> None:
        check(lib().bkma_adv_plan_run(self._h, _ptr(data), nsteps, _stream(stream)))

    def describe(self):
        cfg = TunedConfig()
        buf = ctypes.create_string_buffer(128)
        check(lib().bkma_adv_plan_describe(self._h, ctypes.byref(cfg), buf, 128))
        return cfg, buf.value.decode()

    def close(self) -> None:
        if self._h:
            lib().bkma_adv_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class AdvBatch:
    """A ragged batch of advection problems (bkma_adv_batch_*).  `arrays` are
    contiguous device tensors, `params` their AdvParams."""

    def __init__(self, arrays, params, dtype, autotune: bool = False, stream=None, config=None):
        """config = (lds_budget, threads, steps_per_pass) builds the batch with that configuration
        (bkma_adv_batch_create_with) instead of the heuristic or autotuned one."""
        n = len(arrays)
        probs = (AdvProblem * n)()
        for j, (a, p) in enumerate(zip(arrays, params)):
            probs[j].data = _ptr(a).value
            probs[j].params = p
        self.arrays = list(arrays)
        self._h = ctypes.c_void_p()
        if config is not None:
            budget, threads, ns = config
            check(lib().bkma_adv_batch_create_with(probs, n, dtype, budget, threads, ns, _stream(stream),
                                                   ctypes.byref(self._h)))
        else:
            check(lib().bkma_adv_batch_create(probs, n, dtype, int(autotune), _stream(stream),
                                              ctypes.byref(self._h)))

    def run(self, nsteps: int, graph: bool = True, stream=None) -> None:
        check(lib().bkma_adv_batch_run(self._h, nsteps, int(graph), _stream(stream)))

    def run_loop(self, nsteps: int, mode: int, stream=None) -> None:
        check(lib().bkma_adv_batch_run_loop(self._h, nsteps, mode, _stream(stream)))

    def describe(self):
        v = [ctypes.c_uint64() for _ in range(4)]
        g = ctypes.c_uint32()
        check(lib().bkma_adv_batch_describe(self._h, *[ctypes.byref(x) for x in v], ctypes.byref(g)))
        return {"cells": v[0].value, "items": v[1].value, "segmented_items": v[2].value,
                "lds_budget": v[3].value, "grid": g.value,
                "steps_per_pass": lib().bkma_adv_batch_steps_per_pass(self._h)}

    def close(self) -> None:
        if self._h:
            lib().bkma_adv_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Conv1dPlan:
    """Textbook in-place Conv1d (bkma_conv1d_plan_*): x[B][C][L] device tensor,
    weight[C][C][K] and bias[C] device tensors (F64 for CONV_F64, else F32)."""

    def __init__(self, mode, io_dtype, B, C, L, K, weight, bias, stream=None):
        self._h = ctypes.c_void_p()
        check(lib().bkma_conv1d_plan_create(mode, io_dtype, B, C, L, K, dtype_code(weight), _ptr(weight),
                                            _ptr(bias), _stream(stream), ctypes.byref(self._h)))

    def run(self, x, stream=None) -> None:
        check(lib().bkma_conv1d_plan_run(self._h, _ptr(x), _stream(stream)))

    def set_input_scale(self, scale: float) -> None:
        check(lib().bkma_conv1d_plan_set_input_scale(self._h, scale))

    def name(self) -> str:
        buf = ctypes.create_string_buffer(128)
        check(lib().bkma_conv1d_plan_describe(self._h, buf, 128))
        return buf.value.decode()

    def close(self) -> None:
        if self._h:
            lib().bkma_conv1d_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
