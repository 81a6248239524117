"""The sweep's markdown formatter (bkma_amd.sweep_report) on hand-written JSON
lines in the sweep's output format: rows in run order, skipped runs with their
reason, the conv table and the performance-portability summary."""
import json

import bkma_amd.sweep_report as R


def test_report_tables(tmp_path):
    recs = [
        {"kind": "advection", "size": "e5", "n0": 4, "n1": 1024, "n2": 1, "impl": "ndrange", "pref_wg_size": 128,
         "dtype": "f64", "status": "skipped: NDRange: too long"},
        {"kind": "advection", "size": "e5", "n0": 4, "n1": 1024, "n2": 1, "impl": "tuned", "dtype": "f32",
         "kernel": "line_shfl<f32,cells8,wg128,steps256>", "ms_per_step": 0.5, "gcell_s": 3.0, "GBps": 24.0,
         "arch_eff": 0.003, "l1_err": 1.5e-8, "status": "ok", "app_eff": 1.0},
        {"kind": "conv1d", "config": 0, "batch_size": 16, "input_length": 512, "kernel_size": 3, "channels": 1,
         "impl": "adaptivewg", "pref_wg_size": 512, "dtype": "f64", "ms_per_step": 0.06, "gcell_s": 140.1,
         "GBps": 2241.7, "result": 33.854, "batch_invariant": True, "status": "ok", "app_eff": 1.0},
        {"kind": "summary", "advection_pp": {"tuned/f32": {"arch": 0.04, "app": 0.98, "shapes": 1}},
         "conv1d_pp": {"adaptivewg/w512/f64": {"arch": 0.1, "app": 1.0, "shapes": 1}}},
    ]
    f = tmp_path / "sweep.jsonl"
    f.write_text("progress line\n" + "\n".join(json.dumps(r) for r in recs) + "\n")
    out = R.report(str(f), command="python -m bkma_amd.sweep --sizes e5").splitlines()
    assert "    python -m bkma_amd.sweep --sizes e5" in out
    adv = [l for l in out if l.startswith("| e5 (4, 1024, 1)")]
    assert len(adv) == 2
    assert adv[0].endswith("| skipped: NDRange: too long |")
    assert "| tuned |  | f32 | 0.5 | 3.0 | 24.0 | 0.003 | 1.0 | 1.50e-08 | line_shfl<f32,cells8,wg128,steps256> |" in adv[1]
    assert "| (16, 512, 3, 1) | 0.06 | 140.1 | 2241.7 | 33.85 | True |" in out
    assert "| advection tuned/f32 | 0.04 | 0.98 |" in out
    assert "| conv1d adaptivewg/w512/f64 | 0.1 | 1.0 |" in out
