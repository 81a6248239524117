"""Textbook Conv1d plans (bkma_conv1d_plan_*) on the GPU: MFMA operand
layout, FP64 bit-exact against the oracle, FP32 / BF16 / BF16x3 within their
stated tolerances of the FP64 result, and the in-place tail contract."""
import ctypes
import os

import numpy as np
import pytest

import bkma_amd as B
import oracle as O

pytestmark = pytest.mark.gpu

PROBES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "libbkma_probes.so")
_P = None


def _torch():
    import torch
    return torch


def _probe(name, *tensors, n=None):
    """Calls a probe of tests/cpp/libbkma_probes.so (single MFMAs and the kernels' FP8 conversion) on device
    tensors, on torch's current stream."""
    global _P
    torch = _torch()
    if _P is None:
        B.lib()
        if not os.path.exists(PROBES):
            pytest.fail(f"{PROBES} is missing: build it with __graft_entry__.build()")
        _P = ctypes.CDLL(PROBES)
    args = [ctypes.c_void_p(t.data_ptr()) for t in tensors]
    if n is not None:
        args.append(ctypes.c_ulonglong(n))
    args.append(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert getattr(_P, name)(*args) == 0, name


def test_mfma_32x32x16_bf16_layout(cuda):
    # Small integers are exact in bf16 and in the f32 accumulator.
    torch = _torch()
    rng = np.random.default_rng(0)
    A = rng.integers(-4, 5, size=(32, 16)).astype(np.float32)
    Bm = rng.integers(-4, 5, size=(16, 32)).astype(np.float32)
    a = np.zeros((64, 8), np.float32)
    b = np.zeros((64, 8), np.float32)
    for l in range(64):
        for e in range(8):
            a[l, e] = A[l % 32, 8 * (l // 32) + e]   # A[m][k]: m = l%32, k = 8*(l/32) + e
            b[l, e] = Bm[8 * (l // 32) + e, l % 32]  # B[k][n]: n = l%32, k = 8*(l/32) + e
    ta = torch.from_numpy(a).to(torch.bfloat16).cuda()
    tb = torch.from_numpy(b).to(torch.bfloat16).cuda()
    d = torch.zeros((64, 16), dtype=torch.float32, device="cuda")
    _probe("probe_mfma_bf16", ta, tb, d)
    torch.cuda.synchronize()
    D = A @ Bm
    got = np.zeros((32, 32), np.float32)
    dh = d.cpu().numpy()
    for l in range(64):
        for r in range(16):
            got[8 * (r // 4) + 4 * (l // 32) + r % 4, l % 32] = dh[l, r]
    assert np.array_equal(got, D)


def test_dpp_wave_shl_lane_layout(cuda):
    # The tap-reuse conv kernel moves every column of an accumulator down by one lane with DPP
    # wave_shl:1 (dpp_ctrl 0x130): lane l reads lane l+1 across the whole wave, and lane 63, which has
    # no source, gets the old value (0 here).
    torch = _torch()
    a = torch.arange(100, 164, dtype=torch.int32, device="cuda")
    out = torch.empty(64, dtype=torch.int32, device="cuda")
    _probe("probe_dpp_wave_shl", a, out)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    an = a.cpu().numpy()
    assert np.array_equal(o[:63], an[1:])
    assert o[63] == 0


def _case(Bn, C, L, K, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((Bn, C, L))
    w = rng.standard_normal((C, C, K)) / np.sqrt(C * K)
    b = rng.standard_normal(C) * 0.1
    return x, w, b


def _ref(x, w, b):
    r = x.copy()
    O.conv_torch_f64(r, w, b)
    return r


@pytest.mark.parametrize("shape", [(3, 16, 100, 5), (2, 128, 300, 5), (2, 40, 77, 3), (1, 256, 90, 7)])
def test_conv_f64_bit_exact(cuda, shape):
    torch = _torch()
    Bn, C, L, K = shape
    x, w, b = _case(Bn, C, L, K, 1)
    tx = torch.from_numpy(x).cuda()
    plan = B.Conv1dPlan(B.CONV_F64, B.F64, Bn, C, L, K, torch.from_numpy(w).cuda(), torch.from_numpy(b).cuda())
    plan.run(tx)
    torch.cuda.synchronize()
    assert np.array_equal(tx.cpu().numpy(), _ref(x, w, b))


def _rel(got, ref, Lout):
    d = got[:, :, :Lout].astype(np.float64) - ref[:, :, :Lout]
    return np.abs(d).max() / np.abs(ref[:, :, :Lout]).max(), np.linalg.norm(d) / np.linalg.norm(ref[:, :, :Lout])


# bf16 storage rounds every output to 8 bits: rel_max is then bounded by the
# output rounding (~2^-8 of the largest output), whatever the products' precision.
# bf16x3 (two-term split x = xh + xl, w = wh + wl, three products) misses xl*wl
# and the two split remainders, each at most 2^-18 of |x*w|: 3 * 2^-18 = 1.1e-5 of
# every product (test_conv_bf16x3_error_bound checks it per output); measured
# rel_max 4.9e-6 on these data, asserted at 1e-5.  bf16x6 (three-term split, six
# products) represents x and w exactly in FP32 and meets 1e-6.
@pytest.mark.parametrize("mode,io,tol", [(B.CONV_F32, B.F32, 1e-6), (B.CONV_BF16, B.F32, 1e-2),
                                         (B.CONV_BF16X3, B.F32, 1e-5), (B.CONV_BF16, B.BF16, 1e-2),
                                         (B.CONV_BF16X3, B.BF16, 8e-3), (B.CONV_BF16X6, B.F32, 1e-6)],
                         ids=["f32", "bf16_io32", "bf16x3_io32", "bf16_io16", "bf16x3_io16", "bf16x6_io32"])
@pytest.mark.parametrize("K", [5, 3])
@pytest.mark.parametrize("L", [517, 516], ids=["L517", "L516"])  # scalar and vector staging loads
def test_conv_reduced_precision(cuda, mode, io, tol, K, L):
    torch = _torch()
    Bn, C = 3, 128
    if io == B.BF16 and K != 5:
        pytest.skip("bf16 storage kernels are compiled for K = 5")
    x, w, b = _case(Bn, C, L, K, 2)
    tdt = torch.float32 if io == B.F32 else torch.bfloat16
    tx = torch.from_numpy(x).to(tdt).cuda()
    x_in = tx.double().cpu().numpy()  # the values the kernel actually reads
    plan = B.Conv1dPlan(mode, io, Bn, C, L, K, torch.from_numpy(w).float().cuda(),
                        torch.from_numpy(b).float().cuda())
    print(plan.name(), flush=True)
    plan.run(tx)
    torch.cuda.synchronize()
    got = tx.double().cpu().numpy()
    ref = _ref(x_in, w.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64))
    Lout = L - K + 1
    rmax, rl2 = _rel(got, ref, Lout)
    print(f"rel_max {rmax:.3e} rel_L2 {rl2:.3e}", flush=True)
    assert rmax <= tol
    assert np.array_equal(got[:, :, Lout:], x_in[:, :, Lout:])  # tail kept


def test_conv_bf16x3_error_bound(cuda):
    # Per output: |y - y64| <= 3 * 2^-18 * S + n * 2^-24 * S + 2^-24 |y64|, with S = sum_ic,k |w||x| (the
    # split's missing terms, then FP32 accumulation of n = 3 C K products in any order, then the FP32 store).
    torch = _torch()
    Bn, C, L, K = 2, 128, 400, 5
    x, w, b = _case(Bn, C, L, K, 21)
    tx = torch.from_numpy(x).float().cuda()
    x_in = tx.double().cpu().numpy()
    wf = w.astype(np.float32).astype(np.float64)
    bf = b.astype(np.float32).astype(np.float64)
    plan = B.Conv1dPlan(B.CONV_BF16X3, B.F32, Bn, C, L, K, torch.from_numpy(w).float().cuda(),
                        torch.from_numpy(b).float().cuda())
    plan.run(tx)
    torch.cuda.synchronize()
    Lout = L - K + 1
    got = tx.double().cpu().numpy()[:, :, :Lout]
    ref = _ref(x_in, wf, bf)[:, :, :Lout]
    s_abs = np.abs(x_in).copy()
    O.conv_torch_f64(s_abs, np.abs(wf), np.zeros(C))
    S = s_abs[:, :, :Lout]
    bound = (3 * 2.0 ** -18 + 3 * C * K * 2.0 ** -24) * S + 2.0 ** -24 * np.abs(ref)
    err = np.abs(got - ref)
    print(f"bf16x3: max err/bound {np.max(err / bound):.3f}, max err/(2^-18 S) {np.max(err / (2.0 ** -18 * S)):.3f}",
          flush=True)
    assert np.all(err <= bound)


def test_conv_bf16x6_needs_f32_storage(cuda):
    torch = _torch()
    w = torch.zeros((128, 128, 5), dtype=torch.float32, device="cuda")
    b = torch.zeros(128, dtype=torch.float32, device="cuda")
    with pytest.raises(B.BkmaUnsupported):
        B.Conv1dPlan(B.CONV_BF16X6, B.BF16, 2, 128, 100, 5, w, b)


def test_conv_unsupported_shape_is_reported(cuda):
    torch = _torch()
    w = torch.zeros((96, 96, 5), dtype=torch.float32, device="cuda")
    b = torch.zeros(96, dtype=torch.float32, device="cuda")
    with pytest.raises(B.BkmaUnsupported):
        B.Conv1dPlan(B.CONV_BF16, B.F32, 2, 96, 100, 5, w, b)


# ---------------------------------------------------------------------------
# FP8 (OCP e4m3) building blocks of the FP8 kernels
# ---------------------------------------------------------------------------
def test_fp8_conversion_matches_ocp_e4m3(cuda):
    # Round to nearest even on the e4m3 grid, ties included; beyond the largest finite value 448 (and for
    # infinities) the kernels' conversion saturates to +-448; NaN stays NaN.
    import fp8_e4m3 as F8
    torch = _torch()
    rng = np.random.default_rng(5)
    grid = F8.DECODE[np.isfinite(F8.DECODE)]
    mids = (np.sort(grid)[1:] + np.sort(grid)[:-1]) / 2  # exact ties between neighbours
    beyond = np.array([448.0, 449.0, 463.9, 464.0, 465.0, 479.0, 480.0, 1e4, 3e38, np.inf])
    x = np.concatenate([rng.uniform(-448, 448, 4000), rng.standard_normal(4000) * 1e-2, grid, mids,
                        rng.standard_normal(4000), beyond, -beyond, rng.uniform(-1e6, 1e6, 1000)])
    x = x.astype(np.float32)
    x = x[: len(x) // 2 * 2]
    tx = torch.from_numpy(x).cuda()
    out = torch.empty(len(x), dtype=torch.uint8, device="cuda")
    _probe("probe_fp8_convert", tx, out, n=len(x))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    exp = F8.encode_saturating(x.astype(np.float64))
    bad = np.nonzero(got != exp)[0]
    assert bad.size == 0, [(float(x[i]), int(got[i]), int(exp[i])) for i in bad[:8]]
    # NaN inputs give a NaN code
    tn = torch.tensor([np.nan, -np.nan, 1.0, np.nan], dtype=torch.float32, device="cuda")
    on = torch.empty(4, dtype=torch.uint8, device="cuda")
    _probe("probe_fp8_convert", tn, on, n=4)
    torch.cuda.synchronize()
    codes = on.cpu().numpy()
    assert np.isnan(F8.decode(codes[[0, 1, 3]])).all() and codes[2] == F8.encode(np.array([1.0]))[0]


def test_mfma_32x32x64_fp8_layout(cuda):
    # Lane l holds A[m = l%32][k = 32*(l/32) + e] and B[k = 32*(l/32) + e][n = l%32],
    # e = 0..31; D as for the bf16 32x32 MFMA.  Small integers are exact in e4m3
    # and their 64-term sums are exact in FP32.
    import fp8_e4m3 as F8
    torch = _torch()
    rng = np.random.default_rng(6)
    A = rng.integers(-4, 5, (32, 64)).astype(np.float64)
    Bm = rng.integers(-4, 5, (64, 32)).astype(np.float64)
    a = np.empty((64, 32), np.uint8)
    b = np.empty((64, 32), np.uint8)
    for l in range(64):
        for e in range(32):
            a[l, e] = F8.encode(A[l % 32, 32 * (l // 32) + e])
            b[l, e] = F8.encode(Bm[32 * (l // 32) + e, l % 32])
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    d = torch.empty((64, 16), dtype=torch.float32, device="cuda")
    _probe("probe_mfma_fp8", ta, tb, d)
    torch.cuda.synchronize()
    dh = d.cpu().numpy()
    got = np.empty((32, 32))
    for l in range(64):
        for r in range(16):
            got[8 * (r // 4) + 4 * (l // 32) + r % 4, l % 32] = dh[l, r]
    assert np.array_equal(got, A @ Bm)


# ---------------------------------------------------------------------------
# FP8 conv plans
# ---------------------------------------------------------------------------
FP8_SHAPES = [(2, 128, 300, 5), (2, 128, 301, 5), (2, 256, 301, 7), (1, 256, 140, 7), (3, 256, 1024, 7)]


@pytest.mark.parametrize("shape", FP8_SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("xscale", [1.0, 2.0])
def test_conv_fp8_exact_on_representable_inputs(cuda, shape, xscale):
    # Small integers: every weight (scaled by its channel's power of two) and
    # every input (divided by the power-of-two input scale) is exact in e4m3,
    # and the FP32 sums stay below 2^24, so the result equals the FP64 oracle.
    torch = _torch()
    Bn, C, L, K = shape
    rng = np.random.default_rng(7)
    x = rng.integers(-3, 4, (Bn, C, L)).astype(np.float64)
    w = rng.integers(-4, 5, (C, C, K)).astype(np.float64)
    w[:, 0, 0] = 4.0  # per-channel max |w| = 4 -> scale 2^-6
    b = rng.integers(-8, 9, C).astype(np.float64)
    tx = torch.from_numpy(x).float().cuda()
    plan = B.Conv1dPlan(B.CONV_FP8, B.F32, Bn, C, L, K, torch.from_numpy(w).float().cuda(),
                        torch.from_numpy(b).float().cuda())
    plan.set_input_scale(xscale)
    print(plan.name(), flush=True)
    plan.run(tx)
    torch.cuda.synchronize()
    got = tx.double().cpu().numpy()
    ref = _ref(x, w, b)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("io", [B.F32, B.BF16], ids=["io32", "io16"])
def test_conv_fp8_matches_quantized_emulation(cuda, io):
    # N(0,1) data: the kernel must equal the FP64 convolution of the e4m3-
    # quantized operands (OCP rounding, per-output-channel power-of-two weight
    # scales) up to FP32 accumulation; the error against the unquantized FP64
    # result is printed.
    import fp8_e4m3 as F8
    torch = _torch()
    Bn, C, L, K = 2, 256, 400, 7
    x, w, b = _case(Bn, C, L, K, 8)
    w32 = w.astype(np.float32).astype(np.float64)
    b32 = b.astype(np.float32).astype(np.float64)
    tdt = torch.float32 if io == B.F32 else torch.bfloat16
    tx = torch.from_numpy(x).to(tdt).cuda()
    x_in = tx.double().cpu().numpy()
    plan = B.Conv1dPlan(B.CONV_FP8, io, Bn, C, L, K, torch.from_numpy(w).float().cuda(),
                        torch.from_numpy(b).float().cuda())
    plan.run(tx)
    torch.cuda.synchronize()
    got = tx.double().cpu().numpy()
    amax = np.abs(w32).reshape(C, -1).max(axis=1)
    scale = 2.0 ** np.ceil(np.log2(amax / 448.0))
    wq = F8.decode(F8.encode(w32 / scale[:, None, None])) * scale[:, None, None]
    xq = F8.decode(F8.encode(x_in))
    emu = _ref(xq, wq, b32)
    exact = _ref(x_in, w32, b32)
    Lout = L - K + 1
    e_max, e_l2 = _rel(got, emu, Lout)
    q_max, q_l2 = _rel(got, exact, Lout)
    print(f"vs quantized emulation rel_max {e_max:.3e}; vs FP64 rel_max {q_max:.3e} rel_L2 {q_l2:.3e}", flush=True)
    # FP32 accumulation of C*K = 1792 exact products: the rounding error stays
    # below ~1e-4 of the largest output; bf16 storage rounds the outputs.
    assert e_max <= (1e-4 if io == B.F32 else 8e-3)
    assert np.array_equal(got[:, :, Lout:], x_in[:, :, Lout:])
    assert q_l2 < 0.1


# More tiles than resident work-groups: the persistent loop, the prefetch of
# the next tile into the other LDS buffer and the halo snapshots all run.
@pytest.mark.parametrize("mode,io,tol", [(B.CONV_BF16, B.F32, 1e-2), (B.CONV_BF16X3, B.F32, 1e-5),
                                         (B.CONV_BF16, B.BF16, 1e-2), (B.CONV_BF16X6, B.F32, 1e-6)],
                         ids=["bf16_io32", "bf16x3_io32", "bf16_io16", "bf16x6_io32"])
def test_conv_mfma_persistent_many_tiles(cuda, mode, io, tol):
    torch = _torch()
    Bn, C, L, K = 40, 128, 1024, 5
    x, w, b = _case(Bn, C, L, K, 9)
    tdt = torch.float32 if io == B.F32 else torch.bfloat16
    tx = torch.from_numpy(x).to(tdt).cuda()
    x_in = tx.double().cpu().numpy()
    plan = B.Conv1dPlan(mode, io, Bn, C, L, K, torch.from_numpy(w).float().cuda(), torch.from_numpy(b).float().cuda())
    plan.run(tx)
    torch.cuda.synchronize()
    got = tx.double().cpu().numpy()
    ref = _ref(x_in, w.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64))
    Lout = L - K + 1
    rmax, rl2 = _rel(got, ref, Lout)
    print(plan.name(), f"rel_max {rmax:.3e} rel_L2 {rl2:.3e}", flush=True)
    assert rmax <= tol
    assert np.array_equal(got[:, :, Lout:], x_in[:, :, Lout:])


@pytest.mark.parametrize("L", [2048, 2047], ids=["L2048", "L2047"])
def test_conv_bf16_storage_many_tiles(cuda, L):
    """BF16 storage with vector (L % 4 == 0) and scalar staging loads: 64 x 16
    tiles are more than the resident work-groups, so the persistent loop, the
    double-buffered staging and the halo snapshots all run."""
    torch = _torch()
    Bn, C, K = 64, 128, 5
    x, w, b = _case(Bn, C, L, K, 13)
    tx = torch.from_numpy(x).to(torch.bfloat16).cuda()
    x_in = tx.double().cpu().numpy()
    plan = B.Conv1dPlan(B.CONV_BF16, B.BF16, Bn, C, L, K, torch.from_numpy(w).float().cuda(),
                        torch.from_numpy(b).float().cuda())
    plan.run(tx)
    torch.cuda.synchronize()
    got = tx.double().cpu().numpy()
    ref = _ref(x_in, w.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64))
    Lout = L - K + 1
    rmax, rl2 = _rel(got, ref, Lout)
    print(plan.name(), f"rel_max {rmax:.3e} rel_L2 {rl2:.3e}", flush=True)
    assert rmax <= 1e-2
    assert np.array_equal(got[:, :, Lout:], x_in[:, :, Lout:])


def test_conv_bf16_storage_long_chunks(cuda):
    """BF16 storage at the full line length: 96 x 15 tiles of 280 columns over
    the resident work-groups give chunks of several tiles that cross batch
    elements, so the cross-tile pipeline (a tile's last block is combined in
    the next tile's first block) and the snapshots at chunk ends all run.
    Sampled batch elements are checked against the FP64 oracle."""
    torch = _torch()
    Bn, C, L, K = 96, 128, 4096, 5
    x, w, b = _case(Bn, C, L, K, 17)
    tx = torch.from_numpy(x).to(torch.bfloat16).cuda()
    x_in = tx.double().cpu().numpy()
    plan = B.Conv1dPlan(B.CONV_BF16, B.BF16, Bn, C, L, K, torch.from_numpy(w).float().cuda(),
                        torch.from_numpy(b).float().cuda())
    print(plan.name(), flush=True)
    plan.run(tx)
    torch.cuda.synchronize()
    got = tx.double().cpu().numpy()
    Lout = L - K + 1
    idx = [0, 1, 2, 31, 47, 48, 63, 94, 95]
    ref = _ref(x_in[idx], w.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64))
    rmax, rl2 = _rel(got[idx], ref, Lout)
    print(f"rel_max {rmax:.3e} rel_L2 {rl2:.3e}", flush=True)
    assert rmax <= 1e-2
    assert np.array_equal(got[:, :, Lout:], x_in[:, :, Lout:])


@pytest.mark.parametrize("shape", [(40, 128, 1024, 5), (20, 256, 1024, 7)], ids=lambda s: "x".join(map(str, s)))
def test_conv_fp8_persistent_many_tiles_exact(cuda, shape):
    torch = _torch()
    Bn, C, L, K = shape
    rng = np.random.default_rng(10)
    x = rng.integers(-3, 4, (Bn, C, L)).astype(np.float64)
    w = rng.integers(-4, 5, (C, C, K)).astype(np.float64)
    w[:, 0, 0] = 4.0
    b = rng.integers(-8, 9, C).astype(np.float64)
    tx = torch.from_numpy(x).float().cuda()
    plan = B.Conv1dPlan(B.CONV_FP8, B.F32, Bn, C, L, K, torch.from_numpy(w).float().cuda(),
                        torch.from_numpy(b).float().cuda())
    plan.run(tx)
    torch.cuda.synchronize()
    assert np.array_equal(tx.double().cpu().numpy(), _ref(x, w, b))
