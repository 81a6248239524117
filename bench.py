"""Benchmark of the BKMA advection hot path on one MI355X.

Workload (the headline configuration): f[n0=2048][n1=1024][n2=512], advected
along the strided middle axis, x in [0,1) with dx = 1/1024, v(i0) uniform on
[-1, 1), dt = 1e-3, degree-5 stencil, initial f ~ U(0,1) from seed 1.

A "step" is one in-place advection step over the whole array.  The timed
kernel is the tuned FP32-storage kernel (FP64 foot/index/coefficients and
accumulation).  Beside it, the plain FP64 GPU kernel (the reference's
AdaptiveWg executor and create_optim_params policy, pref_wg_size 256) runs the
same number of steps from the same initial state; its output is the FP64
reference for the FP32 error norms.  The CPU column is the oracle's FP64
restatement of the reference algorithm on a bounded sample of whole slabs.

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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n0", type=int, default=N0, help="(testing only) smaller n0")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "adv_column_stream_f32_2048x1024x512.traffic.json"))
    return ap.parse_args()


def cpu_baseline(target_s: float):
    """Oracle FP64 (reference algorithm, OpenMP over lines) on whole slabs of
    the same workload; the sample size is grown until it takes target_s."""
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
    return {"value": round(cells * steps / el / 1e9, 4), "unit": "Gcell-updates/s", "cores": O.num_threads(),
            "kind": "port",
            "sample": f"{steps} FP64 steps over {slabs} of the {N0} i0-slabs ([{slabs}][{N1}][{N2}]), "
                      f"oracle/bkma_oracle.c (OpenMP over lines), {el:.2f} s after one untimed step"}


def load_traffic(path, bytes_alg):
    try:
        with open(path) as f:
            t = json.load(f)
        return t.get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the advection path is "
              f"embarrassingly parallel over i0 slabs and would shard n0 across GPUs with no per-step "
              f"communication, but this build runs and reports a single MI355X.", file=sys.stderr)
        sys.exit(2)

    import torch
    import bkma_amd as B

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
    _, name32 = plan32.describe()
    plan32.run(f32, args.warmup)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    plan32.run(f32, args.steps)
    e1.record(stream)
    torch.cuda.synchronize()
    wall32 = time.perf_counter() - t0
    kern_ms32 = e0.elapsed_time(e1) / args.steps
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

    bytes_alg = cells * 2 * 4  # 2 * sizeof(float) per cell-update
    achieved = bytes_alg / (kern_ms32 * 1e-3) / 1e9
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
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic(args.traffic, bytes_alg),
                     "kernel_avg_ms": round(kern_ms32, 5), "algorithmic_bytes_per_launch": bytes_alg},
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
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
