"""Ragged batches (bkma_adv_batch_*): every problem of a batch advanced by the
persistent, graph-captured launches must equal the oracle's result for that
problem alone, on every graph replay."""
import numpy as np
import pytest

import bkma_amd as B
import oracle as O

pytestmark = pytest.mark.gpu


def _specs(seed, count, n1_max=4096, n0_max=12):
    """Ragged shapes drawn like the benchmark's batch: n1 log-uniform (multiple
    of 64), n0 uniform, n2 in {1, 8, 64}, dt in [5e-4, 2e-3]."""
    out = []
    for j in range(count):
        u = [O.lib().orc_uniform(seed, 4 * j + k) for k in range(4)]
        n1 = max(64, int(round(np.exp(np.log(64) + u[0] * (np.log(n1_max) - np.log(64))) / 64)) * 64)
        n0 = 2 + int(u[1] * (n0_max - 1))
        n2 = (1, 8, 64)[min(2, int(u[2] * 3))]
        dt = 5e-4 + u[3] * 1.5e-3
        out.append((n0, n1, n2, dt))
    return out


def _make(dtype, specs, seed0):
    import torch
    tdt = torch.float64 if dtype == B.F64 else torch.float32
    arrays, params, hosts = [], [], []
    for j, (n0, n1, n2, dt) in enumerate(specs):
        arrays.append(torch.empty((n0, n1, n2), dtype=tdt, device="cuda"))
        params.append(B.make_params(n0, n1, n2, dt=dt, minRealVx=-1, maxRealVx=1))
    return arrays, params


def _fill(arrays, seed0):
    hosts = []
    for j, a in enumerate(arrays):
        B.fill_uniform(a, seed0 + j)
    import torch
    torch.cuda.synchronize()
    return [a.cpu().numpy() for a in arrays]


def _oracle(hosts, specs, dtype, nsteps, ns=1):
    """One batch run of nsteps steps: passes of ns steps, then the rest (FP32: one rounding per pass; FP64
    storage keeps every intermediate step in FP64 anyway)."""
    refs = []
    for h, (n0, n1, n2, dt) in zip(hosts, specs):
        r = h.copy()
        po = O.make_params(n0, n1, n2, dt=dt, minRealVx=-1, maxRealVx=1)
        if dtype == B.F64:
            O.advect_f64(r, po, nsteps)
        else:
            O.advect_f32_blocked(r, po, nsteps, ns)
        refs.append(r)
    return refs


@pytest.mark.parametrize("dtype", [B.F32, B.F64], ids=["f32", "f64"])
@pytest.mark.parametrize("autotune", [False, True], ids=["heuristic", "autotuned"])
def test_batch_graph_replays_bit_exact(cuda, dtype, autotune):
    import torch
    specs = _specs(3, 10)
    print("specs", specs, flush=True)
    arrays, params = _make(dtype, specs, 100)
    batch = B.AdvBatch(arrays, params, dtype, autotune=autotune)  # before the data is set
    info = batch.describe()
    print(info, flush=True)
    refs = _fill(arrays, 100)
    for rep in range(3):
        batch.run(4, graph=True)
        torch.cuda.synchronize()
        # eager work between replays (a plain copy and a norm)
        _ = arrays[0].clone().sum().item()
        refs = _oracle(refs, specs, dtype, 4, info["steps_per_pass"])
        for j, (a, r) in enumerate(zip(arrays, refs)):
            assert np.array_equal(a.cpu().numpy(), r), (rep, j, specs[j])
    batch.close()


# Whole-line items (one with its rows' feet kept across the steps, one without) and chained segments, advanced
# by fused passes; runs of 5 steps (graph) and 3 steps (no graph) end with a shorter pass.
FUSED_SPECS = [(3, 8192, 8, 2e-3), (2, 16384, 1, 1e-3), (4, 640, 64, 1e-3), (5, 100, 1, 2e-3),
               (3, 2048, 1, 1.5e-3)]


@pytest.mark.parametrize("dtype", [B.F32, B.F64], ids=["f32", "f64"])
@pytest.mark.parametrize("ns", [2, 3, 4])
def test_batch_fused_passes_bit_exact(cuda, dtype, ns):
    import torch
    arrays, params = _make(dtype, FUSED_SPECS, 400)
    batch = B.AdvBatch(arrays, params, dtype, config=(98304, 256, ns))
    info = batch.describe()
    print(info, flush=True)
    assert info["steps_per_pass"] == ns and info["segmented_items"] > 0 and info["items"] > info["segmented_items"]
    refs = _fill(arrays, 400)
    for nsteps, graph in ((5, True), (3, False), (5, True)):
        batch.run(nsteps, graph=graph)
        torch.cuda.synchronize()
        refs = _oracle(refs, FUSED_SPECS, dtype, nsteps, ns)
        for j, (a, r) in enumerate(zip(arrays, refs)):
            assert np.array_equal(a.cpu().numpy(), r), (nsteps, j, FUSED_SPECS[j])
    batch.close()


def test_batch_fused_passes_need_256_threads(cuda):
    arrays, params = _make(B.F32, FUSED_SPECS[:2], 500)
    with pytest.raises(B.BkmaError):
        B.AdvBatch(arrays, params, B.F32, config=(98304, 1024, 2))


def test_batch_without_graph_and_segments(cuda):
    import torch
    # long lines force segments at the 32 KB budget of the heuristic-free path;
    # the last problem's shift (~410 cells) exceeds its segment length, so its
    # segments are not chained (one segment per item)
    specs = [(3, 8192, 8, 2e-3), (2, 16384, 1, 1e-3), (4, 640, 64, 1e-3), (5, 100, 1, 2e-3),
             (2, 16384, 8, 2.5e-2)]
    arrays, params = _make(B.F32, specs, 200)
    batch = B.AdvBatch(arrays, params, B.F32, autotune=False)
    info = batch.describe()
    print(info, flush=True)
    assert info["segmented_items"] > 0
    hosts = _fill(arrays, 200)
    batch.run(3, graph=False)
    torch.cuda.synchronize()
    refs = _oracle(hosts, specs, B.F32, 3)
    for a, r in zip(arrays, refs):
        assert np.array_equal(a.cpu().numpy(), r)


@pytest.mark.parametrize("mode", [0, 1], ids=["reference_loop", "tuned_loop"])
def test_batch_loop_baselines(cuda, mode):
    import torch
    specs = _specs(5, 6)
    arrays, params = _make(B.F64, specs, 300)
    batch = B.AdvBatch(arrays, params, B.F64, autotune=False)
    hosts = _fill(arrays, 300)
    batch.run_loop(3, mode)
    torch.cuda.synchronize()
    refs = _oracle(hosts, specs, B.F64, 3)
    for a, r in zip(arrays, refs):
        assert np.array_equal(a.cpu().numpy(), r)
