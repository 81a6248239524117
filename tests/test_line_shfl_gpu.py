"""GPU parity of LINE_SHFL, the contiguous-line (n2 = 1) advection kernel that
keeps each line in registers and takes the stencil neighbours from the
adjacent lanes through DPP wavefront shifts.

FP64 storage is bit-identical to the oracle's FP64 restatement of the
reference for any steps per pass; FP32 storage is bit-identical to the
oracle's restatement of the tuned FP32 arithmetic with the plan's steps per
pass (orc_advect_f32_blocked).  The cases reach every window start the kernel
compiles for 4 and 8 cells per lane (six-cell windows on lines with one
shift, seven-cell windows on lines with a shift pair), every wavefront count
per line (1 to 16, powers of two or not), and a step count that leaves a
shorter last pass."""
import ctypes
import os

import numpy as np
import pytest

import bkma_amd as B
import oracle as O

pytestmark = pytest.mark.gpu

PROBES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "libbkma_probes.so")


def _torch():
    import torch
    return torch


@pytest.mark.parametrize("which,name", [(0, "wave_shl"), (1, "wave_rol"), (2, "wave_shr"), (3, "wave_ror")])
def test_dpp_wave_moves(cuda, which, name):
    # lane l reads lane l+1 (shl), (l+1) mod 64 (rol), l-1 (shr), (l-1) mod 64
    # (ror); a lane with no source lane keeps its old value
    torch = _torch()
    B.lib()
    if not os.path.exists(PROBES):
        pytest.fail(f"{PROBES} is missing: build it with __graft_entry__.build()")
    P = ctypes.CDLL(PROBES)
    a = torch.arange(100, 164, dtype=torch.int32, device="cuda")
    old = torch.arange(1000, 1064, dtype=torch.int32, device="cuda")
    out = torch.empty(64, dtype=torch.int32, device="cuda")
    st = P.probe_dpp_wave_move(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(old.data_ptr()),
                               ctypes.c_void_p(out.data_ptr()), ctypes.c_ulonglong(which),
                               ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == 0
    torch.cuda.synchronize()
    o, an, on = out.cpu().numpy(), a.cpu().numpy(), old.cpu().numpy()
    lanes = np.arange(64)
    if name == "wave_shl":
        exp = np.where(lanes < 63, an[np.minimum(lanes + 1, 63)], on)
    elif name == "wave_rol":
        exp = an[(lanes + 1) % 64]
    elif name == "wave_shr":
        exp = np.where(lanes > 0, an[np.maximum(lanes - 1, 0)], on)
    else:
        exp = an[(lanes - 1) % 64]
    assert np.array_equal(o, exp), (o, exp)


@pytest.mark.parametrize("r", list(range(1, 16)))
def test_dpp_row_ror(cuda, r):
    # row_ror:r: lane l of each 16-lane row reads lane (l - r) mod 16 of the same row, so the cell at
    # offset o of a 16-cell line held one cell per lane arrives through row_ror:((16 - o) mod 16)
    torch = _torch()
    B.lib()
    P = ctypes.CDLL(PROBES)
    a = torch.arange(100, 164, dtype=torch.int32, device="cuda")
    out = torch.empty(64, dtype=torch.int32, device="cuda")
    st = P.probe_dpp_row_ror(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(out.data_ptr()), ctypes.c_ulonglong(r),
                             ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == 0
    torch.cuda.synchronize()
    lanes = np.arange(64)
    exp = a.cpu().numpy()[(lanes & ~15) | ((lanes - r) & 15)]
    assert np.array_equal(out.cpu().numpy(), exp)


def _shfl_config(ns, cells=8):
    c = B.TunedConfig()
    c.strategy = B.STRATEGY_LINE_SHFL
    c.vec = cells
    c.steps_per_pass = ns
    return c


def _line_shift_sets(po, n0, n1):
    """Shifts d = i1 - leftNode of every row of every line (oracle foot)."""
    out = []
    for i0 in range(n0):
        ds = set()
        for i1 in range(n1):
            d = i1 - O.adv_foot(po, i0, i1)[0]
            d = d - n1 if d > n1 // 2 else d + n1 if d < -(n1 // 2) else d
            ds.add(d)
        out.append(tuple(sorted(ds)))
    return out


def _run(host, p, dtype, nsteps, cfg):
    torch = _torch()
    tdt = torch.float64 if dtype == B.F64 else torch.float32
    d = torch.from_numpy(host).to(dtype=tdt, device="cuda")
    plan = B.AdvPlan(p, dtype, config=cfg)
    got_cfg, name = plan.describe()
    print("plan", name, flush=True)
    plan.run(d, nsteps)
    torch.cuda.synchronize()
    plan.close()
    return d.cpu().numpy(), got_cfg, name


# (dt * n1, minRealVx, maxRealVx, n0, minRealX), per cells per lane.  On
# x in [0.1, 1.1) the rounded coordinates make the rows of an integer shift
# s = v round to both sides, so shifts by halves give one shift per line at
# the half-integers and shift pairs at the integers.
#  - 8 cells: s = v from -4.5 to 5.5, one shift d = -4 .. 6 (six-cell windows
#    from -8 to 2) and pairs {-4,-3} .. {5,6} (seven-cell windows from -8 to
#    2): every window start compiled; fractional s = 5.3 v on x in [0, 1).
#  - 4 cells: s = v from -1.5 to 1.5, one shift d = -1 .. 2 (windows from -4
#    to -1) and pairs {-1,0} .. {1,2} (windows from -4 to -2); fractional
#    s = 1.6 v.
SHIFT_CASES = {8: [(1.0, -4.5, 6.0, 21, 0.1), (5.3, -0.93, 1.0, 7, 0.0)],
               4: [(1.0, -1.5, 2.0, 7, 0.1), (1.6, -0.93, 1.0, 7, 0.0)]}
EXPECTED_SETS = {8: ({(d,) for d in range(-4, 7)}, {(d, d + 1) for d in range(-4, 6)}),
                 4: ({(d,) for d in range(-1, 3)}, {(d, d + 1) for d in range(-1, 2)})}
LINES = {8: [512, 1024, 1536, 2048, 4096], 4: [256, 512, 768, 1024, 2048, 3840, 4096]}
BIT_EXACT_CASES = [(cells, k, n1) for cells in (8, 4) for k in (0, 1) for n1 in LINES[cells]]


@pytest.mark.parametrize("ns", [1, 3, 16])
@pytest.mark.parametrize("cells,case,n1", BIT_EXACT_CASES,
                         ids=[f"cells{c}-{'halves' if k == 0 else 'fractional'}-n{n}" for c, k, n in BIT_EXACT_CASES])
@pytest.mark.parametrize("dtype", [B.F64, B.F32], ids=["f64", "f32"])
def test_line_shfl_bit_exact(cuda, dtype, cells, case, n1, ns):
    dtn, vmin, vmax, n0, x0 = SHIFT_CASES[cells][case]
    dt = dtn / n1  # dx = 1/n1
    p = B.make_params(n0, n1, 1, dt=dt, minRealVx=vmin, maxRealVx=vmax, minRealX=x0, maxRealX=x0 + 1)
    po = O.make_params(n0, n1, 1, dt=dt, minRealVx=vmin, maxRealVx=vmax, minRealX=x0, maxRealX=x0 + 1)
    if case == 0 and n1 == LINES[cells][0]:
        sets = _line_shift_sets(po, n0, n1)
        singles, pairs = EXPECTED_SETS[cells]
        assert {s for s in sets if len(s) == 1} == singles
        assert {s for s in sets if len(s) == 2} == pairs
    seed = 11 + n1 + ns + cells
    if dtype == B.F32:
        host = O.uniform(n0 * n1, seed, np.float32).reshape(n0, n1, 1)
    else:
        host = O.uniform(n0 * n1, seed).reshape(n0, n1, 1)
    nsteps = 2 * ns + 1
    got, cfg, name = _run(host, p, dtype, nsteps, _shfl_config(ns, cells))
    assert name.startswith("line_shfl<")
    assert cfg.steps_per_pass == ns and cfg.vec == cells and cfg.waves_per_wg == n1 // (64 * cells)
    ref = host.copy()
    if dtype == B.F32:
        O.advect_f32_blocked(ref, po, nsteps, ns)
    else:
        O.advect_f64(ref, po, nsteps)
    bits = np.uint32 if dtype == B.F32 else np.uint64
    assert np.array_equal(got.view(bits), ref.view(bits)), np.abs(got - ref).max()


@pytest.mark.parametrize("cells", [8, 4])
def test_line_shfl_c0_workload_bit_exact(cuda, cells):
    # [256][1024][1] FP64, sin(2 pi x) + 0.1 U(0,1) (seed 0), 100 steps in passes of 64 and 36
    n0, n1 = 256, 1024
    p = B.make_params(n0, n1, 1, dt=1e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, 1, dt=1e-3, minRealVx=-1, maxRealVx=1)
    x = (np.arange(n1) / n1)[None, :, None]
    host = np.sin(2 * np.pi * x) + 0.1 * O.uniform(n0 * n1, 0).reshape(n0, n1, 1)
    got, _, _ = _run(host, p, B.F64, 100, _shfl_config(64, cells))
    ref = host.copy()
    O.advect_f64(ref, po, 100)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("shape,dtn,vmin,cells", [
    ((4, 1024, 2), 1.0, 0.0, 8),   # lines strided by n2 = 2
    ((4, 1000, 1), 1.0, 0.0, 8),   # n1 not a power of two
    ((4, 8192, 1), 1.0, 0.0, 8),   # 16 wavefronts of 8 cells per line
    ((4, 4608, 1), 1.0, 0.0, 8),   # 9 wavefronts of 8 cells per line
    ((4, 8192, 1), 1.0, 0.0, 4),   # 32 wavefronts of 4 cells per line
    ((4, 128, 1), 1.0, 0.0, 4),    # half a wavefront per line
    ((4, 1024, 1), 1.0, 0.0, 6),   # 6 cells per lane
    ((2, 1024, 1), 7.5, 1.0, 8),   # shifts 7.5 and 9.375: window starts -10 and -12
    ((2, 1024, 1), 2.5, 1.0, 4),   # shifts 2.5 and 3.125: window starts -5 and -6
    ((2, 1024, 1), 2.0, -1.0, 4),  # shift -2: window start 0, past the lane g+1
], ids=["strided", "ragged_n1", "n1_8192", "n1_4608", "n1_8192_cells4", "n1_128", "cells6", "large_shift",
        "large_shift_cells4", "negative_shift_cells4"])
def test_line_shfl_rejections(cuda, shape, dtn, vmin, cells):
    n0, n1, n2 = shape
    p = B.make_params(n0, n1, n2, dt=dtn / n1, minRealVx=vmin, maxRealVx=vmin + 0.5)
    with pytest.raises(B.BkmaUnsupported):
        B.AdvPlan(p, B.F64, config=_shfl_config(4, cells))


def test_line_shfl_steps_per_pass_bounds(cuda):
    p = B.make_params(4, 1024, 1, dt=1e-3, minRealVx=-1, maxRealVx=1)
    with pytest.raises(B.BkmaUnsupported):
        B.AdvPlan(p, B.F64, config=_shfl_config(257))
    plan = B.AdvPlan(p, B.F64, config=_shfl_config(256))
    plan.close()


@pytest.mark.parametrize("dtype", [B.F64, B.F32], ids=["f64", "f32"])
def test_autotune_contiguous_lines(cuda, dtype):
    # the autotuner times LINE_SHFL beside LINE_LDS and the plain executors;
    # whichever it picks stays bit-exact
    n0, n1 = 64, 2048
    p = B.make_params(n0, n1, 1, dt=1e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, 1, dt=1e-3, minRealVx=-1, maxRealVx=1)
    npdt = np.float64 if dtype == B.F64 else np.float32
    host = O.uniform(n0 * n1, 21, npdt).reshape(n0, n1, 1)
    torch = _torch()
    d = torch.from_numpy(host).to(device="cuda")
    plan = B.AdvPlan(p, dtype, autotune=True)
    cfg, name = plan.describe()
    print("plan", name, flush=True)
    plan.run(d, 7)
    torch.cuda.synchronize()
    plan.close()
    got = d.cpu().numpy()
    ref = host.copy()
    if dtype == B.F32:
        O.advect_f32_blocked(ref, po, 7, max(cfg.steps_per_pass, 1))
    else:
        O.advect_f64(ref, po, 7)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("dtype", [B.F64, B.F32], ids=["f64", "f32"])
def test_heuristic_plan_for_contiguous_lines(cuda, dtype):
    # without the autotuner, contiguous lines go to LINE_SHFL where it fits and
    # to LINE_LDS where it does not, never to the column kernels, whose 64
    # lanes per wavefront span the n2 axis
    for n1, prefix in ((1024, "line_shfl<"), (1000, "line_lds<")):
        p = B.make_params(64, n1, 1, dt=1e-3, minRealVx=-1, maxRealVx=1)
        plan = B.AdvPlan(p, dtype)
        _, name = plan.describe()
        plan.close()
        assert name.startswith(prefix), (n1, name)


# Lines of 16 cells, one per 16-lane DPP row: (dt * n1, minRealVx, dv, n0, minRealX); v = minRealVx + i0 * dv.
#  - the reference's e3 velocity law on 37 lines (two full 16-line groups and a partial one): s = 0.016 v,
#    first-tap offsets -3 and -2;
#  - s = v by halves on x in [0.1, 1.1) (integer shifts give shift pairs): offsets -1..0, -6..-4, -4..-2
#    (a window of 8 cells), and 0 alone.
ROW_CASES = [(0.016, -1.0, 2.0 / 37, 37, 0.0), (1.0, -2.5, 0.5, 3, 0.1), (1.0, 2.0, 0.5, 3, 0.1),
             (1.0, -0.5, 0.5, 4, 0.1), (1.0, -2.5, 0.5, 1, 0.0)]


@pytest.mark.parametrize("ns", [1, 3, 16])
@pytest.mark.parametrize("case", ROW_CASES, ids=["e3_law", "offsets-1_0", "offsets-6_-4", "offsets-4_-2", "offset0"])
@pytest.mark.parametrize("dtype", [B.F64, B.F32], ids=["f64", "f32"])
def test_line_shfl_rows_bit_exact(cuda, dtype, case, ns):
    dtn, vmin, dv, n0, x0 = case
    n1 = 16
    kw = dict(dt=dtn / n1, minRealVx=vmin, maxRealVx=vmin + dv * n0, minRealX=x0, maxRealX=x0 + 1)
    p = B.make_params(n0, n1, 1, **kw)
    po = O.make_params(n0, n1, 1, **kw)
    seed = 41 + n0 + ns
    if dtype == B.F32:
        host = O.uniform(n0 * n1, seed, np.float32).reshape(n0, n1, 1)
    else:
        host = O.uniform(n0 * n1, seed).reshape(n0, n1, 1)
    nsteps = 2 * ns + 1
    got, cfg, name = _run(host, p, dtype, nsteps, _shfl_config(ns, 1))
    assert name.startswith("line_shfl<") and ",row16," in name
    ref = host.copy()
    if dtype == B.F32:
        O.advect_f32_blocked(ref, po, nsteps, ns)
    else:
        O.advect_f64(ref, po, nsteps)
    bits = np.uint32 if dtype == B.F32 else np.uint64
    assert np.array_equal(got.view(bits), ref.view(bits)), np.abs(got - ref).max()


@pytest.mark.parametrize("n1,vmin,dv,n0", [(16, -1.5, 1.0, 4), (32, -0.5, 0.5, 3), (16, 3.5, 0.5, 2)],
                         ids=["offsets_spread_3", "n1_32", "shift_beyond_window"])
def test_line_shfl_rows_rejections(cuda, n1, vmin, dv, n0):
    p = B.make_params(n0, n1, 1, dt=1.0 / n1, minRealVx=vmin, maxRealVx=vmin + dv * n0, minRealX=0.1, maxRealX=1.1)
    with pytest.raises(B.BkmaUnsupported):
        B.AdvPlan(p, B.F64, config=_shfl_config(4, 1))
