"""Textbook Conv1d plans (bkma_conv1d_plan_*) on the GPU: MFMA operand
layout, FP64 bit-exact against the oracle, FP32 / BF16 / BF16x3 within their
stated tolerances of the FP64 result, and the in-place tail contract."""
import numpy as np
import pytest

import bkma_amd as B
import oracle as O

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


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
    B.mfma_layout_probe(ta, tb, d)
    torch.cuda.synchronize()
    D = A @ Bm
    got = np.zeros((32, 32), np.float32)
    dh = d.cpu().numpy()
    for l in range(64):
        for r in range(16):
            got[8 * (r // 4) + 4 * (l // 32) + r % 4, l % 32] = dh[l, r]
    assert np.array_equal(got, D)


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
@pytest.mark.parametrize("mode,io,tol", [(B.CONV_F32, B.F32, 1e-6), (B.CONV_BF16, B.F32, 1e-2),
                                         (B.CONV_BF16X3, B.F32, 1e-4), (B.CONV_BF16, B.BF16, 1e-2),
                                         (B.CONV_BF16X3, B.BF16, 8e-3)],
                         ids=["f32", "bf16_io32", "bf16x3_io32", "bf16_io16", "bf16x3_io16"])
@pytest.mark.parametrize("K", [5, 3])
def test_conv_reduced_precision(cuda, mode, io, tol, K):
    torch = _torch()
    Bn, C, L = 3, 128, 517
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


def test_conv_unsupported_shape_is_reported(cuda):
    torch = _torch()
    w = torch.zeros((96, 96, 5), dtype=torch.float32, device="cuda")
    b = torch.zeros(96, dtype=torch.float32, device="cuda")
    with pytest.raises(B.BkmaUnsupported):
        B.Conv1dPlan(B.CONV_BF16, B.F32, 2, 96, 100, 5, w, b)
