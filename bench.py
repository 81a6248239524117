"""Benchmark of the BKMA advection / Conv1d hot path on one MI355X.

Headline workload (the line `value` and `roofline` describe): f[n0=2048][n1=1024][n2=512], advected along the
strided middle axis, x in [0,1) with dx = 1/1024, v(i0) uniform on [-1, 1), dt = 1e-3, degree-5 stencil,
initial f ~ U(0,1) from seed 1.  A "step" is one in-place advection step over the whole array.  The timed kernel
is the tuned FP32-storage kernel (FP64 foot/index/coefficients and accumulation); a plan may fuse several steps
into one pass over HBM (one launch), keeping the intermediate steps in FP64.  Beside it, the plain FP64 GPU
kernel (the reference's AdaptiveWg executor and create_optim_params policy, pref_wg_size 256) runs the same
number of steps from the same initial state; its output is the FP64 reference for the FP32 error norms.

`roofline` is the headline kernel's use of HBM: the bytes one launch must move (each FP32 cell read and written
once, whatever the number of steps it fuses) over its average duration, against 8 TB/s.  The per-step
equivalent (8 B per cell-update) is reported as `effective_GBps`, and the FP64 vector issue of the stencil sums
as `fp64_valu`.

The other workloads run by default too, each with its own `roofline`:
  c0  plain FP64 [256][1024][1], 100 steps (cache-resident: timing only);
  c2  Conv1d x[1024][128][4096], w[128][128][5]: FP64, FP32, bf16 MFMA (FP32 and BF16 storage), bf16x3,
      bf16x6, with errors against the FP64 plan;
  c3  64 ragged problems, FP32, 200 steps in one hipGraph, against loops of per-problem launches;
  c4  long lines f[256][262144][4] FP64/FP32, 20 steps; FP8-e4m3 Conv1d x[512][256][8192], w[256][256][7].
The CPU column is the oracle's FP64 restatement of the reference algorithm on a bounded sample of whole slabs
of the headline workload.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "parallel-advection_amd"))

N0, N1, N2 = 2048, 1024, 512
DT = 1e-3
SEED = 1
HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK_TFLOPS = 78.6   # MI355X vector FP64
FP32_VALU_PEAK_TFLOPS = 157.3  # MI355X vector FP32
BF16_DENSE_PEAK_TFLOPS = 2500.0
FP8_DENSE_PEAK_TFLOPS = 5000.0
FLOP_PER_CELL_STEP = 12  # the reference's six multiplies and six adds per cell-update
ALL_CONFIGS = ("c0", "c1", "c2", "c3", "c4")


def parse(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n0", type=int, default=N0, help="(testing only) smaller n0 of the headline workload")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--configs", default=",".join(ALL_CONFIGS),
                    help="comma list of the workloads to run besides the headline one (c1 always runs)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "adv_column_pair_f32_2048x1024x512.traffic.txt"),
                    help="JSON summary of the rocprofv3 FETCH_SIZE/WRITE_SIZE passes (hbm_bytes_per_launch)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (3 launches), for scaling rocprofv3 byte counters")
    args = ap.parse_args(argv)
    if args.steps < 1 or args.warmup < 0:
        ap.error("--steps must be at least 1 and --warmup at least 0")
    cfgs = [c for c in args.configs.split(",") if c]
    bad = [c for c in cfgs if c not in ALL_CONFIGS]
    if bad:
        ap.error(f"unknown --configs entries {bad}; known: {','.join(ALL_CONFIGS)}")
    args.config_list = cfgs
    return args


def roofline(bound, work, ms, peak, unit, **extra):
    """Roofline object of one kernel: `work` (bytes or flop per launch) over `ms` (average launch duration)
    against `peak` (GB/s or TFLOP/s)."""
    rate = work / (ms * 1e-3) / (1e9 if unit == "GB/s" else 1e12)
    out = {"bound": bound, "achieved": round(rate, 2), "peak": peak, "unit": unit, "frac": round(rate / peak, 4)}
    out.update(extra)
    return out


def host_cpu():
    """CPU model and the number of cores this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        nproc = os.cpu_count()
    return model, nproc


def cpu_baseline(target_s: float):
    """Oracle FP64 (reference algorithm, OpenMP over lines) on whole slabs of the headline workload; the
    number of steps is sized so the timed part takes about target_s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    slabs = 256
    po = O.make_params(slabs, N1, N2, dt=DT, minRealVx=-1, maxRealVx=1)
    po.dvx = 2.0 / N0  # velocity grid of the full n0 = 2048 problem
    data = O.uniform(slabs * N1 * N2, SEED).reshape(slabs, N1, N2)
    t0 = time.perf_counter()
    O.advect_f64(data, po, 1)
    first = time.perf_counter() - t0
    steps = max(1, min(50, int(target_s / max(first, 1e-3))))
    t0 = time.perf_counter()
    O.advect_f64(data, po, steps)
    el = time.perf_counter() - t0
    cells = slabs * N1 * N2
    model, nproc = host_cpu()
    return {"value": round(cells * steps / el / 1e9, 4), "unit": "Gcell-updates/s", "cores": O.num_threads(),
            "kind": "port", "cpu_model": model, "nproc": nproc,
            "sample": f"{steps} FP64 steps over {slabs} of the {N0} i0-slabs ([{slabs}][{N1}][{N2}]), "
                      f"oracle/bkma_oracle.c (OpenMP over lines, {O.num_threads()} threads), {el:.2f} s after "
                      f"one untimed step"}


def load_traffic(path, plan_name):
    """HBM bytes per launch from the committed rocprofv3 counter summary, if it was taken for the same plan
    (else None)."""
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t.get("hbm_bytes_per_launch") if t.get("plan") == plan_name else None


def log(msg):
    """Progress line on stderr (stdout carries only the JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def event_ms(torch, fn, reps=1):
    """Average duration of fn() over reps calls, from HIP events on torch's current stream (the stream the
    library enqueues on)."""
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s = torch.cuda.current_stream()
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the advection path is "
              f"embarrassingly parallel over i0 slabs and would shard n0 across GPUs with no per-step "
              f"communication, but bench.py runs and reports a single MI355X.", file=sys.stderr)
        sys.exit(2)

    import torch
    import bkma_amd as B

    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4, "launches": 3}))
        return

    n0 = args.n0
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    p = B.make_params(n0, N1, N2, dt=DT, minRealVx=-1, maxRealVx=1)
    if n0 != N0:
        p.dvx = 2.0 / N0
    cells = n0 * N1 * N2

    # ---- initial state, identical for both precisions -------------------
    f64 = torch.empty((n0, N1, N2), dtype=torch.float64, device=dev)
    B.fill_uniform(f64, SEED)
    f32 = f64.to(torch.float32)
    torch.cuda.synchronize()

    # ---- tuned FP32 (the measured path) ---------------------------------
    plan32 = B.AdvPlan(p, B.F32, autotune=not args.no_autotune)
    cfg32, name32 = plan32.describe()
    spp = max(1, cfg32.steps_per_pass)
    plan32.run(f32, args.warmup)
    torch.cuda.synchronize()
    # The timed steps are the plan's passes (full passes of spp steps, then the rest), launched one by one on
    # the current stream with an event before each: the events give every launch's duration.
    passes = [spp] * (args.steps // spp) + ([args.steps % spp] if args.steps % spp else [])
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(passes) + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, k in enumerate(passes):
        ev[i].record(stream)
        plan32.run(f32, k)
    ev[-1].record(stream)
    torch.cuda.synchronize()
    wall32 = time.perf_counter() - t0
    launch_ms = [ev[i].elapsed_time(ev[i + 1]) for i in range(len(passes))]
    # the roofline object describes the longest launch of the run: a full pass of spp steps (fewer only when
    # --steps is below spp)
    spp_l = max(passes)
    full_ms = [t for t, k in zip(launch_ms, passes) if k == spp_l]
    kern_ms32 = sum(full_ms) / len(full_ms)  # average duration of that launch
    gcell32 = cells * args.steps / wall32 / 1e9

    # ---- plain FP64 GPU (reference executor and policy) ----------------
    o = B.create_optim_params(n0, N1, N2, pref_wg_size=256, elem_size=8)
    for _ in range(args.warmup):
        B.adv_run(B.ADAPTIVE_WG, f64, p, o)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        B.adv_run(B.ADAPTIVE_WG, f64, p, o)
    torch.cuda.synchronize()
    wall64 = time.perf_counter() - t0
    gcell64 = cells * args.steps / wall64 / 1e9

    # ---- errors of FP32 vs the FP64 reference after the same steps ------
    rel_max, rel_l2 = B.rel_errors(f32, f64)

    # ---- tuned FP64 (bit-identical to plain; reported for completeness) -
    g64 = torch.empty_like(f64)
    B.fill_uniform(g64, SEED)
    plan64 = B.AdvPlan(p, B.F64, autotune=not args.no_autotune)
    _, name64 = plan64.describe()
    plan64.run(g64, args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan64.run(g64, args.steps)
    torch.cuda.synchronize()
    wallt64 = time.perf_counter() - t0
    gcellt64 = cells * args.steps / wallt64 / 1e9
    same64 = bool(torch.equal(g64, f64))
    del g64
    plan32.close()
    plan64.close()

    # the bytes one launch must move: each FP32 cell read and written once
    bytes_launch = cells * 2 * 4
    # the per-step-equivalent rate: 8 B per cell-update, spp_l cell-updates of every cell per launch
    effective_gbs = cells * spp_l * 2 * 4 / (kern_ms32 * 1e-3) / 1e9
    # the same launch against the vector FP64 peak (the stencil sums run in FP64)
    fp64_tflops = cells * spp_l * FLOP_PER_CELL_STEP / (kern_ms32 * 1e-3) / 1e12
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_seconds)

    out = {
        "metric": "Gcell-updates/s + %HBM roofline, tuned FP32 vs plain FP64 GPU at stated err",
        "value": round(gcell32, 2),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall32 / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "storage_dtype": "f32",
        "data": "synthetic: U(0,1) from the counter-based generator, seed 1",
        "config": {"workload": f"strided advection f[{n0}][{N1}][{N2}] along n1, FP32 storage, "
                               f"v in [-1,1), dt=1e-3, degree-5 stencil, in place",
                   "n0": n0, "n1": N1, "n2": N2, "dt": DT, "kernel": name32},
        "roofline": roofline("hbm", bytes_launch, kern_ms32, HBM_PEAK_GBS, "GB/s",
                             traffic=load_traffic(args.traffic, name32) if n0 == N0 else None,
                             kernel=name32, kernel_avg_ms=round(kern_ms32, 5), bytes_per_launch=bytes_launch,
                             steps_per_launch=spp_l, effective_GBps=round(effective_gbs, 1),
                             fp64_valu={"achieved_TFLOPs": round(fp64_tflops, 2), "peak": FP64_VALU_PEAK_TFLOPS,
                                        "frac": round(fp64_tflops / FP64_VALU_PEAK_TFLOPS, 4),
                                        "flop_per_cell_step": FLOP_PER_CELL_STEP}),
        "plain_fp64_gpu": {"value": round(gcell64, 2), "unit": "Gcell-updates/s",
                           "ms_per_step": round(wall64 / args.steps * 1e3, 4),
                           "GBps_alg": round(gcell64 * 16, 1),
                           "optim": {"w": [o.w0, o.w1, o.w2], "g": [o.wg_dispatch.g0, o.wg_dispatch.g2]}},
        "tuned_fp64_gpu": {"value": round(gcellt64, 2), "unit": "Gcell-updates/s",
                           "GBps_alg": round(gcellt64 * 16, 1), "kernel": name64,
                           "bit_identical_to_plain": same64},
        "speedup_fp32_vs_plain_fp64": round(gcell32 / gcell64, 3),
        "error_fp32_vs_fp64": {"rel_max": rel_max, "rel_L2": rel_l2,
                               "after_steps": args.warmup + args.steps},
        "cpu_baseline": cpu,
    }
    del f64, f32
    torch.cuda.empty_cache()
    log(f"c1 done: {out['value']} Gcell/s, kernel {kern_ms32:.3f} ms per launch")
    cfgs = args.config_list
    for key, cfg, fn in (("c0_small_fp64", "c0", measure_c0), ("c2_conv1d", "c2", measure_conv),
                         ("c3_ragged_batch", "c3", measure_batch), ("c4_long_lines", "c4", measure_long),
                         ("c4_conv1d_fp8", "c4", measure_conv_fp8)):
        if cfg in cfgs:
            t0 = time.perf_counter()
            out[key] = fn(torch, B)
            log(f"{key} done in {time.perf_counter() - t0:.1f} s")
    print(json.dumps(out), flush=True)


def measure_c0(torch, B):
    """Plain FP64 advection f[256][1024][1] (100 steps, dt = 1e-3, v in [-1,1), f = sin(2 pi x) + 0.1 U(0,1)
    from seed 0): the reference's AdaptiveWg executor and the tuned FP64 plan, whose results must be
    bit-identical.  0.5 MB per step stays in cache: timing only, no roofline claim."""
    n0, n1, n2, steps = 256, 1024, 1, 100
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    x = (torch.arange(n1, dtype=torch.float64, device="cuda") / n1).reshape(1, n1, 1)
    noise = torch.empty((n0, n1, n2), dtype=torch.float64, device="cuda")
    B.fill_uniform(noise, 0)
    a = torch.sin(2 * torch.pi * x) + 0.1 * noise
    b = a.clone()
    o = B.create_optim_params(n0, n1, n2, pref_wg_size=512, elem_size=8)
    ms_plain = event_ms(torch, lambda: B.adv_run(B.ADAPTIVE_WG, a, p, o), steps)
    plan = B.AdvPlan(p, B.F64, autotune=True)
    _, name = plan.describe()
    ms_tuned = event_ms(torch, lambda: plan.run(b, steps)) / steps
    plan.close()
    cells = n0 * n1 * n2
    return {"workload": f"f[{n0}][{n1}][{n2}] FP64, {steps} steps", "cells": cells,
            "plain_adaptivewg": {"ms_per_step": round(ms_plain, 4), "Gcell_s": round(cells / ms_plain / 1e6, 2)},
            "tuned_fp64": {"kernel": name, "ms_per_step": round(ms_tuned, 4),
                           "Gcell_s": round(cells / ms_tuned / 1e6, 2)},
            "bit_identical": bool(torch.equal(a, b))}


def measure_conv(torch, B):
    """Conv1d x[1024][128][4096], w[128][128][5] (x ~ N(0,1) seed 2, w ~ N(0, 1/(C K))): HIP-event time per
    call and TFLOP/s of each mode; errors of each mode against the FP64 plan (bit-identical to the oracle) on
    the values that mode read."""
    Bn, C, L, K = 1024, 128, 4096, 5
    flop = 2.0 * Bn * C * C * K * (L - K + 1)
    w = torch.empty((C, C, K), dtype=torch.float64, device="cuda")
    B.fill_normal(w, 21, 1.0 / (C * K) ** 0.5)
    b = torch.zeros(C, dtype=torch.float64, device="cuda")
    x64 = torch.empty((Bn, C, L), dtype=torch.float64, device="cuda")
    B.fill_normal(x64, 2)
    x32 = x64.float()
    ref = x64.clone()
    p64 = B.Conv1dPlan(B.CONV_F64, B.F64, Bn, C, L, K, w, b)
    ms64 = event_ms(torch, lambda: p64.run(ref))
    res = {"workload": f"x[{Bn}][{C}][{L}] w[{C}][{C}][{K}] in place", "flop_per_call": flop,
           "f64": {"kernel": p64.name(), "ms": round(ms64, 3), "TFLOPs": round(flop / ms64 / 1e9, 3),
                   "roofline": roofline("valu", flop, ms64, FP64_VALU_PEAK_TFLOPS, "TFLOP/s")}}
    del x64, ref
    torch.cuda.empty_cache()
    wf, bf = w.float(), b.float()
    Lout = L - K + 1
    for name, mode, io in (("f32", B.CONV_F32, B.F32), ("bf16_io_f32", B.CONV_BF16, B.F32),
                           ("bf16_io_bf16", B.CONV_BF16, B.BF16), ("bf16x3_io_f32", B.CONV_BF16X3, B.F32),
                           ("bf16x3_io_bf16", B.CONV_BF16X3, B.BF16), ("bf16x6_io_f32", B.CONV_BF16X6, B.F32)):
        try:
            plan = B.Conv1dPlan(mode, io, Bn, C, L, K, wf, bf)
        except B.BkmaError as e:
            res[name] = {"error": str(e)}
            continue
        xin = x32 if io == B.F32 else x32.to(torch.bfloat16)
        x = xin.clone()
        plan.run(x)  # warm-up and accuracy run
        # error vs FP64 on the values this mode read (bf16 storage rounds x first)
        r2 = xin.double()
        p64.run(r2)
        torch.cuda.synchronize()
        got = x[:, :, :Lout]
        exp = r2[:, :, :Lout]
        mx, mr, s2, r2n = B.error_norms(got.contiguous(), exp.contiguous())
        rel_max, rel_l2 = mx / mr, (s2 / r2n) ** 0.5
        del r2, got, exp
        torch.cuda.empty_cache()
        ms = event_ms(torch, lambda: plan.run(x), 5)
        if mode == B.CONV_F32:
            rl = roofline("valu", flop, ms, FP32_VALU_PEAK_TFLOPS, "TFLOP/s")
        else:
            # bf16x3 and bf16x6 execute 3 and 6 MFMAs per product term: the fraction counts the useful flop
            rl = roofline("mfma", flop, ms, BF16_DENSE_PEAK_TFLOPS, "TFLOP/s",
                          mfma_per_term={B.CONV_BF16: 1, B.CONV_BF16X3: 3, B.CONV_BF16X6: 6}[mode])
        res[name] = {"kernel": plan.name(), "ms": round(ms, 4), "TFLOPs": round(flop / ms / 1e9, 2),
                     "roofline": rl, "rel_max": rel_max, "rel_L2": rel_l2,
                     "hbm_min_bytes": Bn * C * L * x.element_size() + Bn * C * Lout * x.element_size()}
        plan.close()
        del x
        torch.cuda.empty_cache()
    p64.close()
    return res


def measure_batch(torch, B):
    """64 ragged problems (seed 3): n1 log-uniform in [64, 16384] (multiple of 64), n0 uniform in [16, 1024],
    n2 in {1, 8, 64}, dt in [5e-4, 2e-3], FP32, 200 steps; persistent graph-captured batch vs loops of
    per-problem launches."""
    specs = ragged_specs(3, 64)
    arrays = [torch.empty((n0, n1, n2), dtype=torch.float32, device="cuda") for n0, n1, n2, _ in specs]
    params = [B.make_params(n0, n1, n2, dt=dt, minRealVx=-1, maxRealVx=1) for n0, n1, n2, dt in specs]
    batch = B.AdvBatch(arrays, params, B.F32, autotune=True)
    info = batch.describe()
    for j, a in enumerate(arrays):
        B.fill_uniform(a, 3000 + j)
    steps = 200
    cells = info["cells"]
    out = {"problems": len(specs), "cells": cells, "steps": steps, **info}

    batch.run(steps, graph=True)  # capture + first replay
    ms = event_ms(torch, lambda: batch.run(steps, graph=True))
    step_ms = ms / steps
    ns = max(1, info["steps_per_pass"])
    out["graph_persistent"] = {
        "s": round(ms * 1e-3, 4), "Gcell_s": round(cells * steps / ms / 1e6, 2),
        "roofline": roofline("hbm", cells * 8 / ns, step_ms, HBM_PEAK_GBS, "GB/s", ms_per_step=round(step_ms, 5),
                             steps_per_pass=ns, effective_GBps=round(cells * 8 / step_ms / 1e6, 1),
                             note="one pass of steps_per_pass steps = halo snapshot + persistent launch; each "
                                  "FP32 cell is read and written once per pass (achieved: those bytes per step "
                                  "time; effective_GBps: one read and write per step)")}
    te = event_ms(torch, lambda: batch.run(steps, graph=False))
    out["persistent_no_graph"] = {"s": round(te * 1e-3, 4), "Gcell_s": round(cells * steps / te / 1e6, 2)}
    batch.run_loop(1, 1)
    tl = event_ms(torch, lambda: batch.run_loop(steps, 1))
    out["loop_tuned_per_problem"] = {"s": round(tl * 1e-3, 4), "Gcell_s": round(cells * steps / tl / 1e6, 2)}
    steps_ref = 20  # the reference loop is slow: 20 steps, rate scaled
    tr = event_ms(torch, lambda: batch.run_loop(steps_ref, 0))
    out["loop_reference_adaptivewg"] = {"s_per_step": round(tr * 1e-3 / steps_ref, 5),
                                        "Gcell_s": round(cells * steps_ref / tr / 1e6, 2), "steps": steps_ref}
    out["speedup_vs_reference_loop"] = round((tr / steps_ref) / step_ms, 2)
    out["speedup_vs_tuned_loop"] = round(tl / ms, 3)
    batch.close()
    return out


def _mix64(z):
    m = (1 << 64) - 1
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def _uniform(seed, idx):
    """U[0,1) keyed by (seed, idx): the library's counter-based generator (SplitMix64 finalizer over a Weyl
    sequence, 53 random bits)."""
    m = (1 << 64) - 1
    key = _mix64(seed ^ 0x5851F42D4C957F2D)
    return (_mix64((key + (idx + 1) * 0x9E3779B97F4A7C15) & m) >> 11) * 2.0 ** -53


def ragged_specs(seed, count, n1_lo=64, n1_hi=16384, n0_lo=16, n0_hi=1024):
    """Shapes of the ragged batch: problem j uses draws 4j..4j+3 of `seed`."""
    import math
    out = []
    for j in range(count):
        u = [_uniform(seed, 4 * j + k) for k in range(4)]
        n1 = int(round(math.exp(math.log(n1_lo) + u[0] * (math.log(n1_hi) - math.log(n1_lo))) / 64)) * 64
        n1 = max(64, min(n1_hi, n1))
        n0 = n0_lo + int(u[1] * (n0_hi - n0_lo + 1))
        n2 = (1, 8, 64)[min(2, int(u[2] * 3))]
        dt = 5e-4 + u[3] * 1.5e-3
        out.append((n0, n1, n2, dt))
    return out


def measure_conv_fp8(torch, B):
    """Conv1d x[512][256][8192], w[256][256][7], x and w ~ N(0,1): FP8 e4m3 MFMA (per-output-channel weight
    scales, unit input scale, FP32 accumulation) with FP32 and BF16 storage; errors against the FP64 plan on
    the same inputs."""
    Bn, C, L, K = 512, 256, 8192, 7
    flop = 2.0 * Bn * C * C * K * (L - K + 1)
    w = torch.empty((C, C, K), dtype=torch.float64, device="cuda")
    B.fill_normal(w, 41, 1.0)
    b = torch.zeros(C, dtype=torch.float64, device="cuda")
    x64 = torch.empty((Bn, C, L), dtype=torch.float64, device="cuda")
    B.fill_normal(x64, 4)
    x32 = x64.float()
    del x64
    torch.cuda.empty_cache()
    res = {"workload": f"x[{Bn}][{C}][{L}] w[{C}][{C}][{K}] in place, N(0,1)", "flop_per_call": flop}
    wf, bf = w.float(), b.float()
    Lout = L - K + 1
    p64 = B.Conv1dPlan(B.CONV_F64, B.F64, Bn, C, L, K, w, b)
    for name, io in (("fp8_io_f32", B.F32), ("fp8_io_bf16", B.BF16)):
        plan = B.Conv1dPlan(B.CONV_FP8, io, Bn, C, L, K, wf, bf)
        xin = x32 if io == B.F32 else x32.to(torch.bfloat16)
        x = xin.clone()
        plan.run(x)
        ref = xin.double()
        p64.run(ref)
        torch.cuda.synchronize()
        mx, mr, s2, r2n = B.error_norms(x[:, :, :Lout].contiguous(), ref[:, :, :Lout].contiguous())
        rel_max, rel_l2 = mx / mr, (s2 / r2n) ** 0.5
        del ref
        torch.cuda.empty_cache()
        ms = event_ms(torch, lambda: plan.run(x), 5)
        res[name] = {"kernel": plan.name(), "ms": round(ms, 4), "TFLOPs": round(flop / ms / 1e9, 2),
                     "roofline": roofline("mfma", flop, ms, FP8_DENSE_PEAK_TFLOPS, "TFLOP/s"),
                     "rel_max": rel_max, "rel_L2": rel_l2,
                     "hbm_min_bytes": Bn * C * L * x.element_size() + Bn * C * Lout * x.element_size()}
        plan.close()
        del x
        torch.cuda.empty_cache()
    p64.close()
    return res


def measure_long(torch, B):
    """Long-line advection f[256][262144][4] (lines of 1 MiB FP32 / 2 MiB FP64, beyond LDS), U(0,1) seed 4,
    20 steps, v in [-1,1), dt = 1e-3 (shift up to 262 cells): autotuned FP64 and FP32 plans; the FP32 error
    against FP64; the reference's AdaptiveWg executor with its global scratch."""
    n0, n1, n2, steps = 256, 262144, 4, 20
    cells = n0 * n1 * n2
    p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
    out = {"workload": f"f[{n0}][{n1}][{n2}] along n1, {steps} steps", "cells": cells}
    f64 = torch.empty((n0, n1, n2), dtype=torch.float64, device="cuda")
    B.fill_uniform(f64, 4)
    f32 = f64.float()
    for key, arr, dt in (("fp64", f64, B.F64), ("fp32", f32, B.F32)):
        plan = B.AdvPlan(p, dt, autotune=True)
        cfg, name = plan.describe()
        spp = max(1, cfg.steps_per_pass)
        ms = event_ms(torch, lambda: plan.run(arr, steps))
        passes = -(-steps // spp)
        esz = arr.element_size()
        out[key] = {"kernel": name, "s": round(ms * 1e-3, 4), "Gcell_s": round(cells * steps / ms / 1e6, 2),
                    "GBps_alg": round(cells * steps * 2 * esz / ms / 1e6, 1),
                    "roofline": roofline("hbm", cells * 2 * esz, ms / passes, HBM_PEAK_GBS, "GB/s",
                                         steps_per_launch=spp,
                                         note="time per pass includes the halo snapshot launch")}
        plan.close()
    rel_max, rel_l2 = B.rel_errors(f32, f64)
    out["error_fp32_vs_fp64"] = {"rel_max": rel_max, "rel_L2": rel_l2}
    try:
        g = torch.empty_like(f64)
        B.fill_uniform(g, 4)
        o = B.create_optim_params(n0, n1, n2, pref_wg_size=256, elem_size=8)
        scratch, ext = None, None
        if o.mem_space == B.MEM_GLOBAL:
            ext = B.scratch_extents(B.ADAPTIVE_WG, o, n0, n1, n2, 1)  # AdvectionSolver::window() == 1
            scratch = torch.empty(ext[0] * ext[1] * ext[2], dtype=torch.float64, device="cuda")
        ms = event_ms(torch, lambda: B.adv_run(B.ADAPTIVE_WG, g, p, o, scratch=scratch, scratch_extents=ext), 2)
        out["plain_fp64_gpu"] = {"s_per_step": round(ms * 1e-3, 5), "Gcell_s": round(cells / ms / 1e6, 2),
                                 "mem_space": "global" if o.mem_space == B.MEM_GLOBAL else "local"}
        del g, scratch
    except Exception as e:  # the reference's policy may not cover this shape
        out["plain_fp64_gpu"] = {"error": str(e)}
    del f64, f32
    torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    main()
