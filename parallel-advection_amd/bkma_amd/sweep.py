"""Sweep harness over the reference's benchmark shapes, on one GPU.

    python -m bkma_amd.sweep [--sizes e0,e5] [--steps 3] [--skip-conv] [--fp32]
                             [--tuned-only] [--no-autotune]

Advection (benchmark/advection-bench.cpp:18-96, shapes e0..e9 of
benchmark/bench_config.hpp:31-42): FP64 data filled by fill_buffer_adv,
dt = 1e-3, x in [0, 1), v in [-1, 1) (bench_utils.hpp:29-49).  Per shape it
times the reference's two executors, ndrange and adaptivewg, with the
create_optim_params policy at pref_wg_size 128 and 1024 (the benchmark's
argument grid, advection-bench.cpp:110-117), and the tuned plan (FP64,
bit-identical to adaptivewg; with --fp32 also the FP32-storage plan).  Each
run reports ms per step, Gcell/s, GB/s at 2*sizeof(T) per cell-update
(advection-bench.cpp:59-60) and the analytic L1 error of validate_result_adv
after its steps.  The deltas are computed for the actual shape (the
reference benchmark computes them before it sets n0, n1, n2 and then skips
its error check).

Conv1d (benchmark/conv1d-bench.cpp:18-25, 82-152): the six (batch, length,
kernel, channels) configs with the ConvSolver semantics through the
AdaptiveWg executor, weights 1.5, bias 1.0, data 7.3, pref_wg_size 512; the
result counter is sum_and_normalize_conv over L-K+1 positions.

Output: one JSON line per run, then one summary line with the performance
portability of diag/utils.py:316-455 on this one platform: per
implementation, the harmonic mean over the shapes of its architectural
efficiency (GB/s / 8000) and of its application efficiency (best time of the
shape / its time); an implementation that did not run on some shape scores 0.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import bkma_amd as B

HBM_PEAK_GBS = 8000.0

# benchmark/bench_config.hpp:31-40
EXP_SIZES = {
    "e0": (1 << 17, 1 << 14, 1),
    "e1": (1 << 10, 1 << 11, 1 << 10),
    "e2": (1 << 10, 1 << 14, 1 << 7),
    "e3": (1 << 27, 1 << 4, 1),
    "e4": (1 << 20, 1 << 4, 1 << 7),
    "e5": (1 << 21, 1 << 10, 1),
    "e6": (1 << 14, 1 << 10, 1 << 7),
    "e7": (1, 1 << 10, (1 << 6) + 1),
    "e8": (1, 1 << 10, 1 << 21),
    "e9": (1 << 11, 1 << 10, 1 << 10),
}

# benchmark/conv1d-bench.cpp:18-25: (batch_size, input_length, kernel_size, channels)
CONV_CONFIGS = [(16384, 512, 3, 1), (1024, 512, 3, 1), (32768, 128, 3, 9), (16384, 256, 5, 6),
                (16384, 512, 5, 3), (16384, 1024, 11, 1)]


def _emit(rec: dict) -> None:
    print(json.dumps(rec), flush=True)


def _timed(torch, fn, steps: int, warmup: int) -> float:
    """Seconds per call of fn() over `steps` calls after `warmup` calls."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def _progress(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def advection_runs(torch, name, shape, steps, warmup, fp32, tuned_only=False, autotune=True):
    n0, n1, n2 = shape
    cells = n0 * n1 * n2
    runs = []
    impls = [] if tuned_only else [("ndrange", B.ND_RANGE, 128), ("ndrange", B.ND_RANGE, 1024),
                                   ("adaptivewg", B.ADAPTIVE_WG, 128), ("adaptivewg", B.ADAPTIVE_WG, 1024)]
    data = torch.empty((n0, n1, n2), dtype=torch.float64, device="cuda")
    for impl_name, impl, w in impls:
        rec = {"kind": "advection", "size": name, "n0": n0, "n1": n1, "n2": n2, "impl": impl_name,
               "pref_wg_size": w, "dtype": "f64"}
        _progress(f"{name} {impl_name} pref {w}")
        p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1, pref_wg_size=w)
        try:
            o = B.create_optim_params(n0, n1, n2, pref_wg_size=w, elem_size=8)
            B.fill_adv_sin4(data, p)
            sec = _timed(torch, lambda: B.adv_run(impl, data, p, o), steps, warmup)
        except B.BkmaError as e:
            rec["status"] = f"skipped: {e}"
            runs.append(rec)
            continue
        p.maxIter = steps + warmup
        rec.update(_rates(sec, cells, 8))
        rec["l1_err"] = B.validate_adv_l1(data, p)[0]
        rec["status"] = "ok"
        runs.append(rec)
    for dt_code, tdt, label in ((B.F64, torch.float64, "f64"), (B.F32, torch.float32, "f32")):
        if dt_code == B.F32 and not fp32:
            continue
        rec = {"kind": "advection", "size": name, "n0": n0, "n1": n1, "n2": n2, "impl": "tuned", "dtype": label}
        _progress(f"{name} tuned {label}" + ("" if autotune else " (heuristic plan)"))
        p = B.make_params(n0, n1, n2, dt=1e-3, minRealVx=-1, maxRealVx=1)
        try:
            d = data if dt_code == B.F64 else torch.empty((n0, n1, n2), dtype=tdt, device="cuda")
            B.fill_adv_sin4(d, p)
            plan = B.AdvPlan(p, dt_code, autotune=autotune)
            _, rec["kernel"] = plan.describe()
            spp = max(1, plan.describe()[0].steps_per_pass)
            k = max(steps, spp)  # whole passes
            k -= k % spp
            plan.run(d, warmup)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            plan.run(d, k)
            torch.cuda.synchronize()
            sec = (time.perf_counter() - t0) / k
            plan.close()
        except B.BkmaError as e:
            rec["status"] = f"skipped: {e}"
            runs.append(rec)
            continue
        p.maxIter = k + warmup
        rec.update(_rates(sec, cells, d.element_size()))
        rec["l1_err"] = B.validate_adv_l1(d, p)[0]
        rec["status"] = "ok"
        runs.append(rec)
        if d is not data:
            del d
    del data
    torch.cuda.empty_cache()
    _app_efficiency(runs)
    for rec in runs:  # after the shape, so that each line carries its application efficiency
        _emit(rec)
    return runs


def conv_runs(torch, steps, warmup):
    runs = []
    for idx, (batch, length, k, ch) in enumerate(CONV_CONFIGS):
        n0, n1, n2 = batch, length * ch, 1
        rec = {"kind": "conv1d", "config": idx, "batch_size": batch, "input_length": length, "kernel_size": k,
               "channels": ch, "impl": "adaptivewg", "pref_wg_size": 512, "dtype": "f64"}
        data = torch.empty((n0, n1, n2), dtype=torch.float64, device="cuda")
        w = torch.full((k, ch, ch), 1.5, dtype=torch.float64, device="cuda")
        b = torch.full((ch,), 1.0, dtype=torch.float64, device="cuda")
        try:
            o = B.create_optim_params(n0, n1, n2, pref_wg_size=512, elem_size=8)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

            def call():  # one convolution of the 7.3-filled data (the reference times one)
                B.fill_constant(data, 7.3)
                e0.record()
                B.conv_run(B.ADAPTIVE_WG, data, w, b, k, ch, length, o)
                e1.record()
                torch.cuda.synchronize()
                return e0.elapsed_time(e1) * 1e-3

            for _ in range(warmup):
                call()
            sec = sum(call() for _ in range(steps)) / steps
        except B.BkmaError as e:
            rec["status"] = f"skipped: {e}"
            runs.append(rec)
            _emit(rec)
            continue
        rec.update(_rates(sec, n0 * n1 * n2, 8))
        nw = length - k + 1  # compute_output_size(length, kernel_size)
        rec["result"] = B.conv_sum_normalize(data, nw)
        rec["batch_invariant"] = B.conv_validate(data, nw)
        rec["status"] = "ok"
        rec["app_eff"] = 1.0  # one implementation per config: each is its config's best
        runs.append(rec)
        _emit(rec)
        del data
    torch.cuda.empty_cache()
    return runs


def _rates(sec: float, items: int, esz: int) -> dict:
    gbs = items * 2 * esz / sec / 1e9
    return {"ms_per_step": round(sec * 1e3, 5), "gcell_s": round(items / sec / 1e9, 3), "GBps": round(gbs, 1),
            "arch_eff": round(gbs / HBM_PEAK_GBS, 4)}


def _app_efficiency(runs) -> None:
    ok = [r for r in runs if r.get("status") == "ok"]
    if not ok:
        return
    best = min(r["ms_per_step"] for r in ok)
    for r in ok:
        r["app_eff"] = round(best / r["ms_per_step"], 4)


def perf_portability(groups) -> dict:
    """diag/utils.py:compute_pp on one platform, harmonic means over the
    shapes: {impl: {"arch": pp, "app": pp}} (0 if it failed on a shape)."""
    keys = {}
    for runs in groups:
        for r in runs:
            key = r["impl"] + (f"/w{r['pref_wg_size']}" if "pref_wg_size" in r else "") + "/" + r["dtype"]
            keys.setdefault(key, []).append(r)
    out = {}
    for key, rs in keys.items():
        if len(rs) < len(groups) or any(r.get("status") != "ok" for r in rs):
            out[key] = {"arch": 0.0, "app": 0.0, "shapes": len(groups)}
            continue
        out[key] = {"arch": round(len(rs) / sum(1.0 / r["arch_eff"] for r in rs), 4),
                    "app": round(len(rs) / sum(1.0 / r["app_eff"] for r in rs), 4), "shapes": len(rs)}
    return out


def main(argv=None) -> int:
    """Command-line entry: advection shapes, then the conv configs, then the summary."""
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", default=",".join(EXP_SIZES), help="comma list of e0..e9")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fp32", action="store_true", help="also time the FP32-storage tuned plan")
    ap.add_argument("--skip-conv", action="store_true")
    ap.add_argument("--tuned-only", action="store_true", help="time only the tuned plans (no reference executors)")
    ap.add_argument("--no-autotune", action="store_true", help="tuned plans from the heuristic choice")
    args = ap.parse_args(argv)
    import torch

    if not torch.cuda.is_available():
        print("the sweep needs a GPU", file=sys.stderr)
        return 2
    adv = []
    for name in args.sizes.split(","):
        if name not in EXP_SIZES:
            print(f"unknown size {name}", file=sys.stderr)
            return 2
        adv.append(advection_runs(torch, name, EXP_SIZES[name], args.steps, args.warmup, args.fp32,
                                  tuned_only=args.tuned_only, autotune=not args.no_autotune))
    conv = [] if args.skip_conv else [conv_runs(torch, args.steps, args.warmup)]
    summary = {"kind": "summary", "hbm_peak_GBps": HBM_PEAK_GBS,
               "advection_pp": perf_portability(adv) if adv else {},
               "conv1d_pp": perf_portability([[r] for r in conv[0]]) if conv else {}}
    _emit(summary)
    return 0


if __name__ == "__main__":
    sys.exit(main())
