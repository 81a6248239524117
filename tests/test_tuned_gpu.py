"""GPU parity of the tuned advection kernels (bkma_adv_plan_*).

FP64 storage must be bit-identical to the oracle's FP64 restatement of the
reference, whatever the number of steps fused per pass; FP32 storage must be
bit-identical to the oracle's restatement of the tuned FP32 arithmetic (FP64
foot/index/coefficients, fused FP64 accumulation, FP64 steps inside a pass,
one FP32 rounding per pass) and within 1e-6 relative max-norm of the FP64
result."""
import numpy as np
import pytest

import bkma_amd as B
import oracle as O

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _run(host, params, dtype, nsteps, config=None, autotune=False):
    """Runs a plan; returns (result, steps per pass of the plan)."""
    torch = _torch()
    tdt = torch.float64 if dtype == B.F64 else torch.float32
    d = torch.from_numpy(host).to(dtype=tdt, device="cuda")
    plan = B.AdvPlan(params, dtype, autotune=autotune, config=config)
    cfg, name = plan.describe()
    print("plan", name, flush=True)
    plan.run(d, nsteps)
    torch.cuda.synchronize()
    plan.close()
    return d.cpu().numpy(), max(cfg.steps_per_pass, 1)


SHAPES = [(4, 1024, 512), (3, 333, 64), (5, 200, 192), (2, 96, 96), (7, 1000, 130)]


@pytest.mark.parametrize("shape", SHAPES)
def test_tuned_f64_bit_exact(cuda, shape):
    n0, n1, n2 = shape
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    host = O.uniform(n0 * n1 * n2, 1).reshape(shape)
    got, _ = _run(host, p, B.F64, 9)
    ref = host.copy()
    O.advect_f64(ref, po, 9)
    assert np.array_equal(got, ref), np.abs(got - ref).max()


@pytest.mark.parametrize("shape", SHAPES)
def test_tuned_f32_bit_exact_vs_restatement(cuda, shape):
    n0, n1, n2 = shape
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    host = O.uniform(n0 * n1 * n2, 2, np.float32).reshape(shape)
    got, spp = _run(host, p, B.F32, 9)
    ref = host.copy()
    O.advect_f32_blocked(ref, po, 9, spp)
    assert np.array_equal(got, ref), np.abs(got - ref).max()


def _variant_id(c):
    return (f"vec{c.vec}-smax{c.smax}-p{c.prefetch}-wpb{c.waves_per_wg}{'-nt' if c.nontemporal else ''}"
            f"{'-w64' if c.wide_window else ''}-steps{c.steps_per_pass}")


F32_VARIANTS = B.plan_variants(B.F32)
F64_VARIANTS = B.plan_variants(B.F64)


@pytest.mark.parametrize("persistent", [0, 1])
@pytest.mark.parametrize("cfg", F32_VARIANTS, ids=_variant_id)
def test_every_f32_variant(cuda, cfg, persistent):
    n0, n1, n2 = 3, 300, 256
    # |v*dt/dx| <= 0.9 * (smax - 1) so that every variant covers the shift
    dt = 0.9 * max(cfg.smax - 1, 0.25) / n1
    p = B.make_params(n0, n1, n2, dt=dt, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=dt, minRealVx=-1, maxRealVx=1)
    host = O.uniform(n0 * n1 * n2, 3, np.float32).reshape(n0, n1, n2)
    cfg.persistent = persistent
    got, spp = _run(host, p, B.F32, 5, config=cfg)
    assert spp == cfg.steps_per_pass
    ref = host.copy()
    O.advect_f32_blocked(ref, po, 5, spp)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("cfg", F64_VARIANTS, ids=_variant_id)
def test_every_f64_variant(cuda, cfg):
    n0, n1, n2 = 3, 257, 130 if cfg.vec == 1 else 128
    dt = 0.9 * max(cfg.smax - 1, 0.25) / n1
    p = B.make_params(n0, n1, n2, dt=dt, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=dt, minRealVx=-1, maxRealVx=1)
    host = O.uniform(n0 * n1 * n2, 4).reshape(n0, n1, n2)
    got, _ = _run(host, p, B.F64, 5, config=cfg)
    ref = host.copy()
    O.advect_f64(ref, po, 5)
    assert np.array_equal(got, ref)


PAIR_VARIANTS = [(B.F32, c) for c in F32_VARIANTS if c.strategy == B.STRATEGY_COLUMN_PAIR] + \
    [(B.F64, c) for c in F64_VARIANTS if c.strategy == B.STRATEGY_COLUMN_PAIR]


def _slab_shift_sets(po, n0, n1):
    """Shifts d = i1 - leftNode of every row of every slab (oracle foot)."""
    out = []
    for i0 in range(n0):
        ds = set()
        for i1 in range(n1):
            d = i1 - O.adv_foot(po, i0, i1)[0]
            d = d - n1 if d > n1 // 2 else d + n1 if d < -(n1 // 2) else d
            ds.add(d)
        out.append(tuple(sorted(ds)))
    return out


# (dt * n1, minRealVx, maxRealVx, n0): fractional shifts s = 2.8 v (one
# shift per slab, d = ceil(s) in -2..2), and shifts s = v in -3.5..2.5 by
# halves: the half-integers give one shift per slab (-3..3), the integers
# rows that round to both sides (shift pairs {-3,-2} .. {2,3}).
PAIR_SHIFT_CASES = [(2.8, -1.0, 0.8, 9), (1.0, -3.5, 3.0, 13)]


@pytest.mark.parametrize("case", PAIR_SHIFT_CASES, ids=["fractional", "integer"])
@pytest.mark.parametrize("dtype,cfg", PAIR_VARIANTS,
                         ids=[("f32-" if d == B.F32 else "f64-") + _variant_id(c) for d, c in PAIR_VARIANTS])
def test_pair_every_shift_pair(cuda, dtype, cfg, case):
    """Column-pair kernels: the cases reach every compiled six-tap (one
    shift, D in -3..3) and seven-tap (shifts {D, D+1}, D in -3..2) instance;
    results are bit-identical to the oracle."""
    dtn, vmin, vmax, n0 = case
    n1 = 300
    n2 = 64 * cfg.vec
    dt = dtn / n1  # dx = 1/n1
    p = B.make_params(n0, n1, n2, dt=dt, minRealVx=vmin, maxRealVx=vmax)
    po = O.make_params(n0, n1, n2, dt=dt, minRealVx=vmin, maxRealVx=vmax)
    sets = _slab_shift_sets(po, n0, n1)
    if dtn == 2.8:
        assert {s for s in sets if len(s) == 1} == {(-2,), (-1,), (0,), (1,), (2,)}
    else:
        assert {s for s in sets if len(s) == 1} == {(d,) for d in range(-3, 4)}
        assert {s for s in sets if len(s) == 2} == {(d, d + 1) for d in range(-3, 3)}
    if dtype == B.F32:
        host = O.uniform(n0 * n1 * n2, 5, np.float32).reshape(n0, n1, n2)
    else:
        host = O.uniform(n0 * n1 * n2, 6).reshape(n0, n1, n2)
    nsteps = 2 * cfg.steps_per_pass + 1
    got, spp = _run(host, p, dtype, nsteps, config=cfg)
    assert spp == cfg.steps_per_pass
    ref = host.copy()
    if dtype == B.F32:
        O.advect_f32_blocked(ref, po, nsteps, spp)
    else:
        O.advect_f64(ref, po, nsteps)
    bits = np.uint32 if dtype == B.F32 else np.uint64
    assert np.array_equal(got.view(bits), ref.view(bits)), np.abs(got - ref).max()


def test_pair_shift_beyond_range_rejected(cuda):
    cfg = next(c for c in F32_VARIANTS if c.strategy == B.STRATEGY_COLUMN_PAIR)
    n1 = 300
    p = B.make_params(2, n1, 256, dt=3.5 / n1, minRealVx=0, maxRealVx=2)  # s = 3.5: d = 4
    with pytest.raises(B.BkmaUnsupported):
        B.AdvPlan(p, B.F32, config=cfg)


def _plain_config(pref):
    c = B.TunedConfig()
    c.strategy = B.STRATEGY_PLAIN
    c.seg_rows = pref
    c.steps_per_pass = 1
    return c


@pytest.mark.parametrize("pref", [128, 1024])
def test_plain_strategy_bit_exact(cuda, pref):
    """PLAIN runs the reference's AdaptiveWg executor inside a plan."""
    n0, n1, n2 = 5, 300, 24
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    host = O.uniform(n0 * n1 * n2, 8).reshape(n0, n1, n2)
    got, spp = _run(host, p, B.F64, 4, config=_plain_config(pref))
    assert spp == 1
    ref = host.copy()
    O.advect_f64(ref, po, 4)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_plain_strategy_is_fp64_only(cuda):
    p = B.make_params(5, 300, 24, dt=1e-3, minRealVx=-1, maxRealVx=1)
    with pytest.raises(B.BkmaUnsupported):
        B.AdvPlan(p, B.F32, config=_plain_config(256))


def _hybrid_config(pref, percent):
    c = B.TunedConfig()
    c.strategy = B.STRATEGY_HYBRID
    c.seg_rows = pref
    c.chunk_segs = percent
    c.steps_per_pass = 1
    return c


HYBRID_CASES = [
    # (n0, n1, n2, pref_wg_size, percent of slabs with the scratch in LDS)
    (7, 300, 24, 256, 50),     # 3 LDS slabs, 4 global-scratch slabs, concurrent
    (7, 300, 24, 128, 30),     # 2 and 5
    (9, 200, 130, 256, 75),    # n2 above pref_wg_size: w2 = pref
    (5, 300, 24, 256, 100),    # every slab in LDS
    (5, 300, 24, 256, 0),      # every slab with the global scratch
    (3, 24000, 2, 256, 0),     # lines beyond the LDS: global scratch only
]


@pytest.mark.parametrize("case", HYBRID_CASES, ids=lambda c: "x".join(map(str, c[:3])) + f"-w{c[3]}-loc{c[4]}")
def test_hybrid_strategy_bit_exact(cuda, case):
    """HYBRID splits the slabs between the AdaptiveWg executor with LDS
    scratch and with a global scratch: bit-identical to the oracle."""
    n0, n1, n2, pref, pct = case
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    host = O.uniform(n0 * n1 * n2, 10).reshape(n0, n1, n2)
    got, spp = _run(host, p, B.F64, 4, config=_hybrid_config(pref, pct))
    assert spp == 1
    ref = host.copy()
    O.advect_f64(ref, po, 4)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_hybrid_rejections(cuda):
    p = B.make_params(5, 300, 24, dt=1e-3, minRealVx=-1, maxRealVx=1)
    with pytest.raises(B.BkmaUnsupported):
        B.AdvPlan(p, B.F32, config=_hybrid_config(256, 50))
    with pytest.raises(B.BkmaError):
        B.AdvPlan(p, B.F64, config=_hybrid_config(256, 101))
    # lines beyond the LDS cannot keep their scratch there
    q = B.make_params(4, 24000, 2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    with pytest.raises(B.BkmaUnsupported):
        B.AdvPlan(q, B.F64, config=_hybrid_config(256, 50))


@pytest.mark.parametrize("dtype", [B.F64, B.F32])
def test_short_lines(cuda, dtype):
    """Lines of 16 rows (the reference benchmark's n1 = 16 shapes): the
    autotuned plan is bit-identical to the oracle."""
    n0, n1, n2 = 64, 16, 8
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    np_dt = np.float64 if dtype == B.F64 else np.float32
    host = O.uniform(n0 * n1 * n2, 9, np_dt).reshape(n0, n1, n2)
    got, spp = _run(host, p, dtype, 6, autotune=True)
    ref = host.copy()
    if dtype == B.F64:
        O.advect_f64(ref, po, 6)
    else:
        O.advect_f32_blocked(ref, po, 6, spp)
    bits = np.uint64 if dtype == B.F64 else np.uint32
    assert np.array_equal(got.view(bits), ref.view(bits))


def test_integer_shift_known_answer(cuda):
    # v*dt/dx integer (dt = 1/512, n1 = 1024, v = -1 + i0/2): circular shift
    n0, n1, n2 = 4, 1024, 128
    p = B.make_params(n0, n1, n2, dt=1 / 512, minRealVx=-1, maxRealVx=1)
    host = O.uniform(n0 * n1 * n2, 5).reshape(n0, n1, n2)
    got, _ = _run(host, p, B.F64, 1)
    for i0 in range(n0):
        s = int(round((-1 + i0 / 2) * 2))
        assert np.array_equal(got[i0], np.roll(host[i0], s, axis=0))


def test_autotune_picks_a_correct_variant(cuda):
    n0, n1, n2 = 16, 512, 512
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    host = O.uniform(n0 * n1 * n2, 6, np.float32).reshape(n0, n1, n2)
    got, spp = _run(host, p, B.F32, 7, autotune=True)
    ref = host.copy()
    O.advect_f32_blocked(ref, po, 7, spp)
    assert np.array_equal(got, ref)


def _c1_near_identity_params(first=992, n0=64):
    """The strided benchmark's line shape [.][1024][512] and its own slabs first .. first+n0-1, whose
    velocities v = -1 + i0/1024 surround v = 0 (slab 1024): shifts v*dt/dx of a few thousandths of a cell,
    where each step changes a value by less than half an FP32 ulp and per-step rounding loses the most.  A CPU
    scan of every slab with the oracle puts the workload's largest FP32 errors at slabs 1023-1026 and
    their mirrors around slabs 24 and 2024 (shifts of about -1 and +1 cells).  v0 and dvx = 1/1024 are exact,
    so these are the benchmark's own velocities."""
    p = B.make_params(n0, 1024, 512, dt=1e-3, minRealVx=-1 + first / 1024, maxRealVx=-1 + (first + n0) / 1024)
    assert p.dvx == 1 / 1024 and p.minRealVx == -1 + first / 1024
    return p


def test_f32_error_budget_vs_f64(cuda):
    # The autotuned FP32 plan, all 50 steps in one call (its passes keep the intermediate steps in FP64 and
    # round once per pass): rel_max <= 1e-6 against the FP64 plan (bit-identical to the reference's FP64).
    torch = _torch()
    p = _c1_near_identity_params()
    d64 = torch.empty((p.n0, p.n1, p.n2), dtype=torch.float64, device="cuda")
    B.fill_uniform(d64, 1)
    d32 = d64.to(torch.float32)
    plan = B.AdvPlan(p, B.F32, autotune=True)
    cfg, name = plan.describe()
    B.AdvPlan(p, B.F64).run(d64, 50)
    plan.run(d32, 50)
    rel_max, rel_l2 = B.rel_errors(d32, d64)
    print(f"{name}: FP32 vs FP64 after 50 steps in one call: rel_max {rel_max:.3e} rel_L2 {rel_l2:.3e}", flush=True)
    assert rel_max <= 1e-6
    assert rel_l2 <= 3e-7


def test_f32_error_of_one_call_per_step(cuda):
    # The reference's time loop calls bkma_run once per step (src/advection.cpp:55-58).  With FP32 storage
    # every call ends in one FP32 rounding, so 50 calls round 50 times: on these near-identity slabs, where
    # updates below half an ulp are lost at every rounding, the error grows by up to 2^-24 (half an ulp of the
    # values in [1, 2) that the interpolation overshoot reaches) per step: 50 * 2^-24 = 2.98e-6 relative to
    # max|f| >= 1.  The oracle gives 1.44e-6 on these slabs (1.8e-6 over the full workload), above the 1e-6
    # that the plan meets with whole runs per call; the results stay bit-identical to the oracle's
    # one-rounding-per-step restatement.
    torch = _torch()
    p = _c1_near_identity_params(1016, 16)
    d64 = torch.empty((p.n0, p.n1, p.n2), dtype=torch.float64, device="cuda")
    B.fill_uniform(d64, 1)
    d32 = d64.to(torch.float32)
    host = d32.cpu().numpy()
    plan = B.AdvPlan(p, B.F32, autotune=True)
    B.AdvPlan(p, B.F64).run(d64, 50)
    for _ in range(50):
        plan.run(d32, 1)
    rel_max, rel_l2 = B.rel_errors(d32, d64)
    print(f"FP32, 50 calls of one step: rel_max {rel_max:.3e} rel_L2 {rel_l2:.3e}", flush=True)
    assert rel_max <= 50 * 2.0 ** -24
    po = O.make_params(p.n0, p.n1, p.n2, dt=1e-3, minRealVx=p.minRealVx, maxRealVx=p.maxRealVx)
    O.advect_f32_blocked(host, po, 50, 1)
    assert np.array_equal(d32.cpu().numpy().view(np.uint32), host.view(np.uint32))


def test_steps_per_pass_beyond_compiled_is_rejected(cuda):
    p = B.make_params(4, 1024, 512, dt=1e-3, minRealVx=-1, maxRealVx=1)
    cfg = F32_VARIANTS[0]
    cfg.steps_per_pass = 64
    with pytest.raises(B.BkmaUnsupported):
        B.AdvPlan(p, B.F32, config=cfg)


def test_unsupported_shapes_are_reported(cuda):
    p = B.make_params(2, 1024, 8, dt=0.6, minRealVx=-1, maxRealVx=1)  # shift beyond half a line
    with pytest.raises(B.BkmaUnsupported):
        B.AdvPlan(p, B.F32)


# ---------------------------------------------------------------------------
# LINE_LDS: slabs (or segments of them, with halo snapshots) staged in LDS
# ---------------------------------------------------------------------------
def _lds_cfg(seg_rows, chunk_segs=1):
    return B.TunedConfig(strategy=B.STRATEGY_LINE_LDS, seg_rows=seg_rows, chunk_segs=chunk_segs)


LDS_CASES = [
    # (n0, n1, n2, dt, seg_rows, chunk_segs)
    (8, 1024, 1, 1e-3, 0, 1),        # whole line, the FP64 parity configuration's shape
    (5, 777, 3, 2e-3, 0, 1),         # odd sizes
    (3, 2048, 8, 4e-3, 0, 1),        # shift ~8 cells
    (3, 4096, 8, 2e-3, 1000, 1),     # same, segmented (ragged last segment)
    (3, 4096, 8, 2e-3, 1000, 2),     # chained segments, ragged last chunk
    (2, 16384, 4, 1e-3, 1024, 1),    # long line, shift ~16 cells, 16 segments
    (2, 16384, 4, 1e-3, 1024, 4),    # 4 chunks of 4 segments
    (2, 16384, 4, 1e-3, 1024, 16),   # one chunk per slab (halo only at the wrap)
    (2, 16384, 4, 1e-3, 1024, 3),    # chunks not dividing the segments
    (4, 2048, 1, 1e-2, 256, 1),      # shift ~20 cells, segments shorter than the halo span
    (4, 2048, 1, 1e-2, 256, 5),      # same, chained (n2 = 1: scalar path)
    (3, 2048, 3, 3e-2, 128, 4),      # shift ~61 cells, odd n2
    (3, 256, 64, 2e-2, 0, 1),        # n2 = 64 with a shift (~5 cells) beyond the register window
]


@pytest.mark.parametrize("case", LDS_CASES,
                         ids=lambda c: "x".join(map(str, c[:3])) + f"-dt{c[3]}-seg{c[4]}-chunk{c[5]}")
@pytest.mark.parametrize("dtype", [B.F64, B.F32], ids=["f64", "f32"])
def test_line_lds_bit_exact(cuda, case, dtype):
    n0, n1, n2, dt, seg, chunk = case
    p = B.make_params(n0, n1, n2, dt=dt, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=dt, minRealVx=-1, maxRealVx=1)
    npdt = np.float64 if dtype == B.F64 else np.float32
    host = O.uniform(n0 * n1 * n2, 7, npdt).reshape(n0, n1, n2)
    got, spp = _run(host, p, dtype, 4, config=_lds_cfg(seg, chunk))
    assert spp == 1
    ref = host.copy()
    if dtype == B.F64:
        O.advect_f64(ref, po, 4)
    else:
        O.advect_f32_mixed(ref, po, 4)
    assert np.array_equal(got, ref), np.abs(got.astype(np.float64) - ref).max()


def _lds_config(ns, seg_rows=0, slabs=1):
    c = B.TunedConfig()
    c.strategy = B.STRATEGY_LINE_LDS
    c.seg_rows = seg_rows
    c.steps_per_pass = ns
    c.chunk_segs = slabs  # whole lines: slabs staged together
    return c


@pytest.mark.parametrize("slabs", [1, 7])
@pytest.mark.parametrize("ns", [1, 2, 3, 4])
@pytest.mark.parametrize("shape", [(3, 200, 8), (2, 512, 1), (4, 64, 5), (300, 16, 1)])
@pytest.mark.parametrize("dtype", [B.F64, B.F32])
def test_line_lds_fused_steps(cuda, dtype, shape, ns, slabs):
    """Whole-line LINE_LDS advancing the staged lines ns times in LDS: FP64
    bit-identical to the reference's steps, FP32 to the restatement with FP64
    intermediate steps (9 steps: full passes and a shorter last one).  With
    slabs = 7, units stage up to 7 consecutive slabs (a partial last unit for
    n0 = 4 and 300)."""
    n0, n1, n2 = shape
    p = B.make_params(n0, n1, n2, dt=2e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=2e-3, minRealVx=-1, maxRealVx=1)
    np_dt = np.float64 if dtype == B.F64 else np.float32
    host = O.uniform(n0 * n1 * n2, 10, np_dt).reshape(shape)
    got, spp = _run(host, p, dtype, 9, config=_lds_config(ns, slabs=slabs))
    assert spp == ns
    ref = host.copy()
    if dtype == B.F64:
        O.advect_f64(ref, po, 9)
    else:
        O.advect_f32_blocked(ref, po, 9, ns)
    bits = np.uint64 if dtype == B.F64 else np.uint32
    assert np.array_equal(got.view(bits), ref.view(bits))


FUSED_SEG_CASES = [
    # (n0, n1, n2, dt, seg_rows, chunk_segs)
    (3, 4096, 8, 2e-3, 500, 1),      # shift ~8 cells, segments with their own snapshots
    (3, 4096, 8, 2e-3, 500, 3),      # chained segments (carried halo), ragged last chunk
    (2, 16384, 4, 1e-3, 1024, 4),    # shift ~16 cells, 4 chunks of 4 segments
    (2, 16384, 4, 1e-3, 1024, 16),   # one chunk per slab: halos only at the wrap
    (4, 2048, 1, 1e-2, 256, 3),      # shift ~20 cells, n2 = 1 (scalar columns)
    (3, 2048, 6, 1e-2, 512, 2),      # n2 = 6 (pairs of columns)
]


def _lds_seg_config(ns, seg_rows, chunk_segs):
    c = _lds_config(ns, seg_rows=seg_rows)
    c.chunk_segs = chunk_segs
    return c


@pytest.mark.parametrize("ns", [2, 3, 4])
@pytest.mark.parametrize("case", FUSED_SEG_CASES,
                         ids=lambda c: "x".join(map(str, c[:3])) + f"-dt{c[3]}-seg{c[4]}-chunk{c[5]}")
@pytest.mark.parametrize("dtype", [B.F64, B.F32], ids=["f64", "f32"])
def test_line_lds_segments_fused_steps(cuda, dtype, case, ns):
    """Segmented LINE_LDS fusing ns steps: each segment stages ns steps'
    halos (carried, read ahead, or from the chunk's snapshot) and computes
    the rows later steps still need (9 steps: full passes and a shorter
    last one)."""
    n0, n1, n2, dt, seg, chunk = case
    p = B.make_params(n0, n1, n2, dt=dt, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=dt, minRealVx=-1, maxRealVx=1)
    np_dt = np.float64 if dtype == B.F64 else np.float32
    host = O.uniform(n0 * n1 * n2, 11, np_dt).reshape(n0, n1, n2)
    got, spp = _run(host, p, dtype, 9, config=_lds_seg_config(ns, seg, chunk))
    assert spp == ns
    ref = host.copy()
    if dtype == B.F64:
        O.advect_f64(ref, po, 9)
    else:
        O.advect_f32_blocked(ref, po, 9, ns)
    bits = np.uint64 if dtype == B.F64 else np.uint32
    assert np.array_equal(got.view(bits), ref.view(bits)), np.abs(got.astype(np.float64) - ref).max()


WG_CASES = [
    # (n0, n1, n2, dt, seg_rows, chunk_segs, steps per pass)
    (3, 4096, 8, 2e-3, 500, 3, 1),   # chained segments
    (3, 4096, 8, 2e-3, 500, 3, 3),   # chained segments, fused steps
    (2, 16384, 4, 1e-3, 1024, 4, 1),
    (3, 200, 8, 2e-3, 0, 1, 1),      # whole lines
    (300, 16, 1, 2e-3, 0, 7, 4),     # whole lines, slab groups, fused steps
]


@pytest.mark.parametrize("wpb", [8, 16])
@pytest.mark.parametrize("case", WG_CASES, ids=lambda c: "x".join(map(str, c[:3])) + f"-seg{c[4]}-chunk{c[5]}-ns{c[6]}")
@pytest.mark.parametrize("dtype", [B.F64, B.F32], ids=["f64", "f32"])
def test_line_lds_work_group_sizes(cuda, dtype, case, wpb):
    """LINE_LDS with 512- and 1024-thread work-groups: same results as the
    oracle (7 steps: full passes and a shorter last one)."""
    n0, n1, n2, dt, seg, chunk, ns = case
    p = B.make_params(n0, n1, n2, dt=dt, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=dt, minRealVx=-1, maxRealVx=1)
    np_dt = np.float64 if dtype == B.F64 else np.float32
    host = O.uniform(n0 * n1 * n2, 12, np_dt).reshape(n0, n1, n2)
    c = _lds_seg_config(ns, seg, chunk)
    c.waves_per_wg = wpb
    got, spp = _run(host, p, dtype, 7, config=c)
    assert spp == ns
    ref = host.copy()
    if dtype == B.F64:
        O.advect_f64(ref, po, 7)
    else:
        O.advect_f32_blocked(ref, po, 7, ns)
    bits = np.uint64 if dtype == B.F64 else np.uint32
    assert np.array_equal(got.view(bits), ref.view(bits))


def test_line_lds_work_group_size_checked(cuda):
    p = B.make_params(3, 200, 8, dt=1e-3, minRealVx=-1, maxRealVx=1)
    c = _lds_config(1)
    c.waves_per_wg = 6
    with pytest.raises(B.BkmaUnsupported):
        B.AdvPlan(p, B.F32, config=c)


def test_line_lds_chained_fused_steps_need_segments_past_the_halo(cuda):
    # v in {-1, -0.5, 0, 0.5}: backward shifts up to ~21 cells need 23 halo
    # rows a step; 4 steps reach 92 rows, beyond 64-row segments, 2 steps fit
    p = B.make_params(4, 2048, 1, dt=2e-2, minRealVx=-1, maxRealVx=1)
    with pytest.raises(B.BkmaUnsupported):
        B.AdvPlan(p, B.F32, config=_lds_seg_config(4, 64, 4))
    B.AdvPlan(p, B.F32, config=_lds_seg_config(2, 64, 4)).close()


def test_chained_segments_shorter_than_halo_rejected(cuda):
    # v in {-1, -0.5, 0, 0.5}: backward shifts up to ~31 cells need 33 halo
    # rows before a segment, more than a 32-row segment can carry to the next
    p = B.make_params(4, 2048, 4, dt=3e-2, minRealVx=-1, maxRealVx=1)
    with pytest.raises(B.BkmaUnsupported):
        B.AdvPlan(p, B.F32, config=_lds_cfg(32, 4))


def test_autotune_small_n2_long_line(cuda):
    # c4-like shape reduced in n0 and n1: FP32, shift ~66 cells
    n0, n1, n2 = 4, 65536, 4
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    host = O.uniform(n0 * n1 * n2, 8, np.float32).reshape(n0, n1, n2)
    got, spp = _run(host, p, B.F32, 2, autotune=True)
    ref = host.copy()
    O.advect_f32_blocked(ref, po, 2, spp)
    assert np.array_equal(got, ref)


def test_tuned_c0_config_bit_exact(cuda):
    # [256][1024][1] FP64, sin(2 pi x) + 0.1 U(0,1) (seed 0), 100 steps: the
    # tuned plan (whole lines in LDS) against the oracle's reference arithmetic
    n0, n1, n2 = 256, 1024, 1
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    x = (np.arange(n1) / n1)[None, :, None]
    host = np.sin(2 * np.pi * x) + 0.1 * O.uniform(n0 * n1 * n2, 0).reshape(n0, n1, n2)
    got, _ = _run(host, p, B.F64, 100)
    ref = host.copy()
    O.advect_f64(ref, po, 100)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("dtype", [B.F64, B.F32], ids=["f64", "f32"])
def test_zero_steps_is_identity(cuda, dtype):
    p = B.make_params(4, 1024, 512, dt=1e-3, minRealVx=-1, maxRealVx=1)
    npdt = np.float64 if dtype == B.F64 else np.float32
    host = O.uniform(4 * 1024 * 512, 9, npdt).reshape(4, 1024, 512)
    got, _ = _run(host, p, dtype, 0)
    assert np.array_equal(got, host)


@pytest.mark.parametrize("axis", ["n0", "n1", "n2"])
def test_empty_extents_are_rejected(cuda, axis):
    """An empty array is an argument error at every advection entry point
    (plans with and without the autotuner, the plain executors), never a
    launch."""
    torch = _torch()
    p = B.make_params(4, 64, 8, dt=1e-3, minRealVx=-1, maxRealVx=1)
    setattr(p, axis, 0)
    d = torch.zeros(16, dtype=torch.float64, device="cuda")
    for autotune in (False, True):
        with pytest.raises(B.BkmaInvalidArgument):
            B.AdvPlan(p, B.F64, autotune=autotune)
    with pytest.raises(B.BkmaInvalidArgument):
        B.AdvPlan(p, B.F32, config=_lds_config(1))
    o = B.create_optim_params(4, 64, 8, pref_wg_size=128, elem_size=8)
    for impl in (B.ADAPTIVE_WG, B.ND_RANGE):
        with pytest.raises(B.BkmaInvalidArgument):
            B.adv_run(impl, d, p, o)
    # a batch holding one empty problem beside a valid one
    q = B.make_params(2, 64, 8, dt=1e-3, minRealVx=-1, maxRealVx=1)
    e = torch.zeros(2 * 64 * 8, dtype=torch.float64, device="cuda")
    with pytest.raises(B.BkmaInvalidArgument):
        B.AdvBatch([e, d], [q, p], B.F64)


@pytest.mark.parametrize("shape", [(1, 32, 1), (1, 64, 64), (2, 40, 7), (1, 1024, 4096)])
@pytest.mark.parametrize("dtype", [B.F64, B.F32], ids=["f64", "f32"])
def test_small_and_single_slab_shapes(cuda, shape, dtype):
    n0, n1, n2 = shape
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    po = O.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    npdt = np.float64 if dtype == B.F64 else np.float32
    host = O.uniform(n0 * n1 * n2, 10, npdt).reshape(shape)
    got, spp = _run(host, p, dtype, 5)
    ref = host.copy()
    if dtype == B.F64:
        O.advect_f64(ref, po, 5)
    else:
        O.advect_f32_blocked(ref, po, 5, spp)
    assert np.array_equal(got, ref)
