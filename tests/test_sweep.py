"""The sweep harness (bkma_amd.sweep) over the reference's benchmark shapes:
its performance-portability arithmetic on the CPU, and a reduced run on the
GPU whose FP64 results must agree across every executor."""
import pytest

import bkma_amd.sweep as S


def test_reference_shapes_and_configs():
    # benchmark/bench_config.hpp:31-40 and conv1d-bench.cpp:18-25
    assert len(S.EXP_SIZES) == 10 and S.EXP_SIZES["e7"] == (1, 1024, 65)
    assert all(n0 * n1 * n2 == 1 << 31 for k, (n0, n1, n2) in S.EXP_SIZES.items() if k not in ("e7",))
    assert len(S.CONV_CONFIGS) == 6 and S.CONV_CONFIGS[5] == (16384, 1024, 11, 1)


def test_perf_portability_is_the_harmonic_mean():
    runs = [[{"impl": "a", "dtype": "f64", "status": "ok", "arch_eff": 0.5, "app_eff": 1.0},
             {"impl": "b", "dtype": "f64", "status": "ok", "arch_eff": 0.25, "app_eff": 0.5}],
            [{"impl": "a", "dtype": "f64", "status": "ok", "arch_eff": 0.25, "app_eff": 0.5},
             {"impl": "b", "dtype": "f64", "status": "skipped: n1 too large", }]]
    pp = S.perf_portability(runs)
    assert pp["a/f64"]["arch"] == pytest.approx(2 / (1 / 0.5 + 1 / 0.25), abs=1e-4)
    assert pp["a/f64"]["app"] == pytest.approx(2 / (1 / 1.0 + 1 / 0.5), abs=1e-4)
    assert pp["b/f64"] == {"arch": 0.0, "app": 0.0, "shapes": 2}  # failed on one shape


def test_app_efficiency_is_best_over_own():
    runs = [{"status": "ok", "ms_per_step": 2.0}, {"status": "ok", "ms_per_step": 4.0}, {"status": "skipped"}]
    S._app_efficiency(runs)
    assert runs[0]["app_eff"] == 1.0 and runs[1]["app_eff"] == 0.5 and "app_eff" not in runs[2]


@pytest.mark.gpu
def test_reduced_sweep_agrees_across_executors(cuda):
    import torch
    runs = S.advection_runs(torch, "small", (4, 256, 64), 2, 1, True)
    ok = {(r["impl"], r.get("pref_wg_size"), r["dtype"]): r for r in runs if r["status"] == "ok"}
    assert len(ok) == 6, runs
    # every FP64 executor is bit-identical after the same 3 steps, so their L1 errors agree
    errs = {r["l1_err"] for k, r in ok.items() if k[2] == "f64" and k[0] != "tuned"}
    assert len(errs) == 1 and next(iter(errs)) < 1e-6
    assert ok[("tuned", None, "f32")]["l1_err"] < 1e-6


@pytest.mark.gpu
def test_reduced_conv_configs(cuda, monkeypatch):
    import torch
    monkeypatch.setattr(S, "CONV_CONFIGS", [(8, 64, 3, 2), (4, 128, 5, 3)])
    runs = S.conv_runs(torch, 1, 1)
    assert [r["status"] for r in runs] == ["ok", "ok"]
    assert all(r["batch_invariant"] for r in runs)
