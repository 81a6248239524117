"""GPU parity of the plain BKMA executors (AdaptiveWg Local/Global, NDRange,
BasicRange) and the reference ConvSolver against the CPU oracle, through the
C ABI.  The plain kernels keep the reference's FP64 arithmetic, so the bar is
bit-exact equality with the oracle (and 1e-12 relative max-norm at least)."""
import numpy as np
import pytest

import bkma_amd as B
import oracle as O

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _c0_init(n0, n1, n2, seed=0):
    """f = sin(2 pi x_i1) + 0.1 U(0,1), U from the shared generator."""
    x = np.arange(n1) / n1
    u = O.uniform(n0 * n1 * n2, seed).reshape(n0, n1, n2)
    return np.sin(2 * np.pi * x)[None, :, None] + 0.1 * u


def _run_gpu(host, impl, params, nsteps, optim=None, dtype=None, scratch_kind=None):
    torch = _torch()
    dt = dtype or torch.float64
    d = torch.from_numpy(host).to(dtype=dt, device="cuda")
    scratch = None
    if scratch_kind == "global":
        e = B.scratch_extents(B.ADAPTIVE_WG, optim, params.n0, params.n1, params.n2, 1)
        scratch = torch.empty(e, dtype=dt, device="cuda")
    elif scratch_kind == "basic":
        scratch = torch.empty((params.n0, params.n1, params.n2), dtype=dt, device="cuda")
    for _ in range(nsteps):
        B.adv_run(impl, d, params, optim, scratch)
    torch.cuda.synchronize()
    return d.cpu().numpy()


def _assert_bit_exact(got, ref):
    diff = np.abs(got.astype(np.float64) - ref)
    rel = diff.max() / np.abs(ref).max()
    print(f"max|diff| = {diff.max():.3e}  rel_max = {rel:.3e}", flush=True)
    assert rel <= 1e-12
    assert np.array_equal(got, ref)


def test_adaptivewg_local_c0_config_bit_exact(cuda):
    # [256][1024][1] FP64, v in [-1, 1), dt = 1e-3, degree-5 stencil, 100 steps
    p = B.make_params(256, 1024, 1, dt=1e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(256, 1024, 1, dt=1e-3, minRealVx=-1, maxRealVx=1)
    host = _c0_init(256, 1024, 1)
    o = B.create_optim_params(256, 1024, 1, pref_wg_size=512, elem_size=8)
    got = _run_gpu(host, B.ADAPTIVE_WG, p, 100, o)
    ref = host.copy()
    O.advect_f64(ref, po, 100)
    _assert_bit_exact(got, ref)


@pytest.mark.parametrize("shape,pref", [((16, 256, 33), 256), ((3, 1000, 70), 128), ((100, 16, 1), 128),
                                        ((5, 2048, 2), 1024)])
def test_adaptivewg_local_strided(cuda, shape, pref):
    n0, n1, n2 = shape
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    host = O.uniform(n0 * n1 * n2, 1).reshape(shape)
    o = B.create_optim_params(n0, n1, n2, pref_wg_size=pref, elem_size=8)
    print("optim", (o.w0, o.w1, o.w2), (o.wg_dispatch.g0, o.wg_dispatch.g2), flush=True)
    got = _run_gpu(host, B.ADAPTIVE_WG, p, 7, o)
    ref = host.copy()
    O.advect_f64(ref, po, 7)
    _assert_bit_exact(got, ref)


def test_adaptivewg_oversized_work_group_is_rejected(cuda):
    # w0 = pref / n1 * n2 = 128 gives a 128*16*4 work-group (bkma_tools.hpp:86):
    # a launch error in the reference, a clear invalid-argument status here.
    torch = _torch()
    p = B.make_params(160, 16, 4)
    o = B.create_optim_params(160, 16, 4, pref_wg_size=512, elem_size=8)
    d = torch.zeros((160, 16, 4), dtype=torch.float64, device="cuda")
    with pytest.raises(B.BkmaInvalidArgument, match="exceeds 1024"):
        B.adv_run(B.ADAPTIVE_WG, d, p, o)


def test_adaptivewg_global_scratch(cuda):
    n0, n1, n2 = 12, 512, 40
    p = B.make_params(n0, n1, n2, dt=2e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=2e-3, minRealVx=-1, maxRealVx=1)
    host = O.uniform(n0 * n1 * n2, 2).reshape(n0, n1, n2)
    o = B.create_optim_params(n0, n1, n2, pref_wg_size=256, elem_size=8)
    o.mem_space = B.MEM_GLOBAL
    got = _run_gpu(host, B.ADAPTIVE_WG, p, 5, o, scratch_kind="global")
    ref = host.copy()
    O.advect_f64(ref, po, 5)
    _assert_bit_exact(got, ref)


def test_adaptivewg_multi_batch(cuda):
    # host batch loop of bkma_run with several (and ragged) batches per dim
    n0, n1, n2 = 10, 256, 9
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    host = O.uniform(n0 * n1 * n2, 3).reshape(n0, n1, n2)
    o = B.create_optim_params(n0, n1, n2, pref_wg_size=64, elem_size=8)
    o.dispatch_d0 = B.lib().bkma_init_1d_blocking(n0, 4)
    o.dispatch_d2 = B.lib().bkma_init_1d_blocking(n2, 4)
    o.w2 = 2
    got = _run_gpu(host, B.ADAPTIVE_WG, p, 3, o)
    ref = host.copy()
    O.advect_f64(ref, po, 3)
    _assert_bit_exact(got, ref)


def test_adaptivewg_batches_not_multiple_of_grid_stride(cuda):
    # Batches of 5 lines in dims 0 and 2 with w0 = w2 = 2 give g = 2 work-groups, a grid stride of 4 lines:
    # one work-item's second trip lies past its batch.  Every line is still stepped exactly once per call.
    n0, n1, n2 = 10, 256, 10
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    host = O.uniform(n0 * n1 * n2, 13).reshape(n0, n1, n2)
    o = B.create_optim_params(n0, n1, n2, pref_wg_size=64, elem_size=8)
    o.dispatch_d0 = B.lib().bkma_init_1d_blocking(n0, 5)
    o.dispatch_d2 = B.lib().bkma_init_1d_blocking(n2, 5)
    o.w0, o.w2 = 2, 2
    assert (o.dispatch_d0.n_batch, o.dispatch_d0.batch_size) == (2, 5)
    got = _run_gpu(host, B.ADAPTIVE_WG, p, 3, o)
    ref = host.copy()
    O.advect_f64(ref, po, 3)
    _assert_bit_exact(got, ref)


@pytest.mark.parametrize("shape", [(64, 1024, 1), (7, 333, 5)])
def test_ndrange(cuda, shape):
    n0, n1, n2 = shape
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    host = O.uniform(n0 * n1 * n2, 4).reshape(shape)
    got = _run_gpu(host, B.ND_RANGE, p, 4)
    ref = host.copy()
    O.advect_f64(ref, po, 4)
    _assert_bit_exact(got, ref)


def test_ndrange_rejects_long_lines(cuda):
    torch = _torch()
    p = B.make_params(2, 2048, 1)
    d = torch.zeros((2, 2048, 1), dtype=torch.float64, device="cuda")
    with pytest.raises(B.BkmaInvalidArgument, match="exceeds the maximum work-group size"):
        B.adv_run(B.ND_RANGE, d, p)


def test_basicrange(cuda):
    n0, n1, n2 = 9, 300, 7
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    host = O.uniform(n0 * n1 * n2, 5).reshape(n0, n1, n2)
    got = _run_gpu(host, B.BASIC_RANGE, p, 4, scratch_kind="basic")
    ref = host.copy()
    O.advect_f64(ref, po, 4)
    _assert_bit_exact(got, ref)


def test_global_scratch_required(cuda):
    torch = _torch()
    p = B.make_params(4, 64, 4)
    d = torch.zeros((4, 64, 4), dtype=torch.float64, device="cuda")
    o = B.create_optim_params(4, 64, 4, pref_wg_size=64, elem_size=8)
    o.mem_space = B.MEM_GLOBAL
    with pytest.raises(B.BkmaInvalidArgument, match="global scratch"):
        B.adv_run(B.ADAPTIVE_WG, d, p, o)
    with pytest.raises(B.BkmaInvalidArgument, match="scratch"):
        B.adv_run(B.BASIC_RANGE, d, p)


def test_plain_f32_storage(cuda):
    torch = _torch()
    n0, n1, n2 = 32, 512, 16
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    host = O.uniform(n0 * n1 * n2, 6, np.float32).reshape(n0, n1, n2)
    o = B.create_optim_params(n0, n1, n2, pref_wg_size=256, elem_size=4)
    got = _run_gpu(host, B.ADAPTIVE_WG, p, 10, o, dtype=torch.float32)
    ref = host.copy()
    O.advect_f32_plain(ref, po, 10)
    assert np.array_equal(got, ref)


def test_reference_l1_criterion_on_gpu(cuda):
    # Impl.* / ValidateNoIteration semantics through the GPU validator
    torch = _torch()
    p = B.make_params(32, 1024, 32, dt=1e-4, maxIter=0)
    d = torch.empty((32, 1024, 32), dtype=torch.float64, device="cuda")
    B.fill_adv_sin4(d, p)
    hi, lo = B.validate_adv_l1(d, p)
    assert hi == 0.0 and lo == 0.0
    o = B.create_optim_params(32, 1024, 32, pref_wg_size=128, elem_size=8)
    for _ in range(100):
        B.adv_run(B.ADAPTIVE_WG, d, p, o)
    p.maxIter = 100
    hi, lo = B.validate_adv_l1(d, p)
    po = O.make_params(32, 1024, 32, dt=1e-4, maxIter=100)
    f = O.fill_adv_sin4(po)
    O.advect_f64(f, po, 100)
    hi_ref, _ = O.l1_error_adv(f, po)
    print("L1 gpu", hi, "oracle", hi_ref, flush=True)
    assert hi < 1e-6
    assert abs(hi - hi_ref) <= 1e-3 * hi_ref


def test_gpu_uniform_matches_oracle_generator(cuda):
    torch = _torch()
    for dt, npdt in ((torch.float64, np.float64), (torch.float32, np.float32)):
        t = torch.empty(100003, dtype=dt, device="cuda")
        B.fill_uniform(t, 42)
        assert np.array_equal(t.cpu().numpy(), O.uniform(100003, 42, npdt))


def test_error_norms(cuda):
    torch = _torch()
    a = torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64, device="cuda")
    r = torch.tensor([1.0, 2.5, 2.0], dtype=torch.float32, device="cuda")
    mx, mr, s2, r2 = B.error_norms(a, r)
    assert (mx, mr) == (1.0, 2.5) and abs(s2 - 1.25) < 1e-15 and abs(r2 - (1 + 6.25 + 4)) < 1e-12


# ---------------------------------------------------------------------------
# ConvSolver through AdaptiveWg (reference semantics)
# ---------------------------------------------------------------------------
def test_conv_literal_constant_known_answer(cuda):
    torch = _torch()
    C, K, L, n0, n2 = 3, 3, 1024, 4, 8
    n1 = C * L
    data = torch.full((n0, n1, n2), 7.3, dtype=torch.float64, device="cuda")
    w = torch.full((K, C, C), 1.5, dtype=torch.float64, device="cuda")
    b = torch.ones(C, dtype=torch.float64, device="cuda")
    o = B.create_optim_params(n0, n1, n2, pref_wg_size=512, elem_size=8)
    before = B.conv_sum_normalize(data, n1)
    assert abs(before - 7.3) < 1e-12
    B.conv_run(B.ADAPTIVE_WG, data, w, b, K, C, L, o)
    nw = L - K + 1
    assert B.conv_validate(data, nw)
    host = data.cpu().numpy()
    for oc in range(C):  # 1 + C*K*1.5*7.3 = 99.55 up to the order of the 9 additions
        np.testing.assert_allclose(host[:, oc * L: oc * L + nw, :], 1 + C * K * 1.5 * 7.3, rtol=1e-14)
    ref = np.full((n0, n1, n2), 7.3)
    O.conv_literal_f64(ref, np.full((K, C, C), 1.5), np.ones(C), K, C, L)
    assert np.array_equal(host, ref)


@pytest.mark.parametrize("n0,n2,C,L,K", [(3, 2, 4, 37, 5), (2, 5, 8, 64, 3), (4, 1, 3, 1024, 3)])
def test_conv_literal_random_bit_exact(cuda, n0, n2, C, L, K):
    torch = _torch()
    rng = np.random.default_rng(n0 * 100 + K)
    host = rng.standard_normal((n0, C * L, n2))
    w = rng.standard_normal((K, C, C))
    b = rng.standard_normal(C)
    d = torch.from_numpy(host).cuda()
    o = B.create_optim_params(n0, C * L, n2, pref_wg_size=256, elem_size=8)
    B.conv_run(B.ADAPTIVE_WG, d, torch.from_numpy(w).cuda(), torch.from_numpy(b).cuda(), K, C, L, o)
    ref = host.copy()
    O.conv_literal_f64(ref, w, b, K, C, L)
    assert np.array_equal(d.cpu().numpy(), ref)
