"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/*.h declares, and the host-side dispatch policy reproduces the
reference's create_optim_params (src/tools/init.hpp:76-102,
src/core/bkma_tools.hpp:34-153) including its quirks."""
import glob
import os
import re
import subprocess

import pytest

import bkma_amd as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(bkma_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = B.lib()
    declared = _declared_functions()
    assert len(declared) >= 16
    out = subprocess.run(["nm", "-D", "--defined-only", B.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = sorted(declared - exported)
    assert not missing, missing
    for name in declared:
        getattr(L, name)
    assert L.bkma_abi_version() == 1


def test_binding_covers_header():
    assert _declared_functions() <= set(B.EXPORTS)


def test_binding_strategy_codes_match_header():
    text = open(os.path.join(ROOT, "include", "bkma_amd.h")).read()
    codes = {m.group(1): int(m.group(2)) for m in re.finditer(r"\bBKMA_(STRATEGY_\w+)\s*=\s*(\d+)", text)}
    assert len(codes) == 6
    for name, value in codes.items():
        assert getattr(B, name) == value, name


def test_update_deltas():
    p = B.make_params(256, 1024, 1, dt=1e-3, minRealVx=-1, maxRealVx=1)
    assert p.dx == 2.0 ** -10 and p.inv_dx == 1024.0 and p.realWidthX == 1.0 and p.dvx == 2 / 256


def test_policy_c0_64k_lds():
    o = B.create_optim_params(256, 1024, 1, pref_wg_size=512, elem_size=8, local_mem_bytes=64 * 1024)
    assert (o.w0, o.w1, o.w2) == (1, 512, 1)
    assert (o.wg_dispatch.g0, o.wg_dispatch.g2) == (256, 1)
    assert o.mem_space == B.MEM_LOCAL
    assert (o.dispatch_d0.n_batch, o.dispatch_d0.batch_size, o.dispatch_d0.last_batch_size) == (1, 256, 256)


def test_policy_strided_160k_lds():
    # n2 >= pref: w = (1, 1, 256); w2*n1 >= 20480 -> w2 = 20, w1 = 256/20 = 12
    o = B.create_optim_params(2048, 1024, 512, pref_wg_size=256, elem_size=8, local_mem_bytes=160 * 1024)
    assert (o.w0, o.w1, o.w2) == (1, 12, 20)
    assert (o.wg_dispatch.g0, o.wg_dispatch.g2) == (2048, 25)


def test_policy_w0_precedence_quirk():
    # n1*n2 < pref: w0 = pref / n1 * n2 = (512 / 16) * 4 = 128 (bkma_tools.hpp:86)
    o = B.create_optim_params(256, 16, 4, pref_wg_size=512, elem_size=8, local_mem_bytes=64 * 1024)
    assert (o.w0, o.w1, o.w2) == (128, 16, 4)
    with pytest.raises(B.BkmaInvalidArgument, match="A single work-group cannot process more"):
        B.create_optim_params(64, 16, 4, pref_wg_size=512, elem_size=8, local_mem_bytes=64 * 1024)


def test_policy_seq_sizes_go_to_batch_slots():
    # create_optim_params passes seq_size0/2 as n_batchs (init.hpp:89-91)
    o = B.create_optim_params(1024, 1024, 1, pref_wg_size=512, seq_size0=4, elem_size=8, local_mem_bytes=65536)
    assert o.wg_dispatch.g0 == 1024 // 4 and o.wg_dispatch.s0 == 1


def test_policy_line_longer_than_lds_falls_back_to_global():
    # the reference divides by zero here (bkma_tools.hpp:102-105)
    o = B.create_optim_params(256, 262144, 4, pref_wg_size=256, elem_size=8, local_mem_bytes=160 * 1024)
    assert o.mem_space == B.MEM_GLOBAL and o.w2 == 1 and o.w0 == 1 and o.w1 >= 1


def test_init_1d_blocking():
    b = B.lib().bkma_init_1d_blocking(10, 4)
    assert (b.n_batch, b.batch_size, b.last_batch_size) == (3, 4, 2)
    b = B.lib().bkma_init_1d_blocking(8, 4)
    assert (b.n_batch, b.batch_size, b.last_batch_size) == (2, 4, 4)
    b = B.lib().bkma_init_1d_blocking(3, 65535)
    assert (b.n_batch, b.batch_size, b.last_batch_size) == (1, 3, 3)


def test_scratch_extents():
    o = B.create_optim_params(2048, 1024, 512, pref_wg_size=256, elem_size=8, local_mem_bytes=160 * 1024)
    assert B.scratch_extents(B.ADAPTIVE_WG, o, 2048, 1024, 512, 1) == (2048, 500, 1024)
    assert B.scratch_extents(B.ADAPTIVE_WG, o, 2048, 1024, 512, 5) == (2048, 500, 1020)
    assert B.scratch_extents(B.BASIC_RANGE, None, 3, 4, 5, 1) == (3, 4, 5)


def test_compute_without_device_is_an_error_not_a_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    p = B.make_params(4, 16, 1)
    import ctypes
    st = B.lib().bkma_adv_run(B.ADAPTIVE_WG, B.F64, ctypes.c_void_p(16), ctypes.byref(p), None,
                              B.Span3d(None, 0, 0, 0), None)
    assert st == B.BKMA_ERR_DEVICE
    assert b"no HIP device" in B.lib().bkma_last_error()
