"""bkma_adv_run_noalloc (global scratch from the library's static device array) against bkma_adv_run with a
caller-allocated scratch, through the C ABI: bit-identical results on consecutive calls, the second with a
smaller array, for the executors that use a global scratch (AdaptiveWg/Global, BasicRange)."""
import pytest

import bkma_amd as B

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _pair(n0, n1, n2, dtype, seed):
    torch = _torch()
    a = torch.empty((n0, n1, n2), dtype=dtype, device="cuda")
    B.fill_uniform(a, seed)
    return a, a.clone()


def _adaptivewg_global(n0, n1, n2, elem):
    o = B.create_optim_params(n0, n1, n2, pref_wg_size=128, elem_size=elem)
    o.mem_space = B.MEM_GLOBAL
    return o


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
def test_noalloc_adaptivewg_global_matches_caller_scratch(cuda, dtype_name):
    torch = _torch()
    dtype = getattr(torch, dtype_name)
    elem = torch.empty(0, dtype=dtype).element_size()
    for (n0, n1, n2), seed in (((64, 1024, 16), 41), ((8, 512, 4), 42)):  # second call: a smaller array
        print(f"noalloc AdaptiveWg/Global {dtype_name} [{n0}][{n1}][{n2}]", flush=True)
        p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
        o = _adaptivewg_global(n0, n1, n2, elem)
        e = B.scratch_extents(B.ADAPTIVE_WG, o, n0, n1, n2, 1)
        assert e[0] * e[1] * e[2] * elem <= B.lib().bkma_static_scratch_bytes()
        a, b = _pair(n0, n1, n2, dtype, seed)
        scratch = torch.empty(e, dtype=dtype, device="cuda")
        for _ in range(2):
            B.adv_run_noalloc(B.ADAPTIVE_WG, a, p, o)
            B.adv_run(B.ADAPTIVE_WG, b, p, o, scratch)
        torch.cuda.synchronize()
        assert torch.equal(a, b)


def test_noalloc_basicrange_matches_caller_scratch(cuda):
    torch = _torch()
    for (n0, n1, n2), seed in (((16, 256, 8), 43), ((4, 128, 2), 44)):
        print(f"noalloc BasicRange float32 [{n0}][{n1}][{n2}]", flush=True)
        p = B.make_params(n0, n1, n2, dt=2e-3, minRealVx=-1, maxRealVx=1)
        a, b = _pair(n0, n1, n2, torch.float32, seed)
        scratch = torch.empty((n0, n1, n2), dtype=torch.float32, device="cuda")
        for _ in range(2):
            B.adv_run_noalloc(B.BASIC_RANGE, a, p)
            B.adv_run(B.BASIC_RANGE, b, p, None, scratch)
        torch.cuda.synchronize()
        assert torch.equal(a, b)


def test_noalloc_local_needs_no_scratch(cuda):
    torch = _torch()
    n0, n1, n2 = 32, 1024, 1
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    o = B.create_optim_params(n0, n1, n2, pref_wg_size=256, elem_size=8)
    assert o.mem_space == B.MEM_LOCAL
    a, b = _pair(n0, n1, n2, torch.float64, 45)
    B.adv_run_noalloc(B.ADAPTIVE_WG, a, p, o)
    B.adv_run(B.ADAPTIVE_WG, b, p, o)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


def test_noalloc_rejects_scratch_beyond_the_static_array(cuda):
    torch = _torch()
    n0, n1, n2 = 2048, 1024, 16  # BasicRange scratch of 2048*1024*16*8 B = 256 MiB
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    a = torch.zeros((n0, n1, n2), dtype=torch.float64, device="cuda")
    with pytest.raises(B.BkmaInvalidArgument, match="exceeds the static scratch"):
        B.adv_run_noalloc(B.BASIC_RANGE, a, p)
    torch.cuda.synchronize()
    assert int(torch.count_nonzero(a)) == 0  # nothing was launched
