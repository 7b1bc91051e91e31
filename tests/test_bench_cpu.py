"""bench.py's roofline object, assembled from kernel times without a GPU.

The kernel-time keys are the ones the timing helpers return (k_step_fused for
the single-pass step, k_flux / k_update for the two-pass step); the profile
summaries that the object reads (traffic, VALU instructions per cell) are
looked up under the kernel's own name, which differs from the pass key for
float domains (k_walk_step, k_walk_flux)."""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


CELLS = 2048 * 2048


def test_single_pass_float_names_the_column_walk_and_reads_its_profile(bench):
    per_kernel, r = bench.roofline_fields({"k_step_fused": 0.070, "k_finalize+gaps": 0.006}, 0.076, 4, False,
                                          CELLS, "rainfall_runoff_2048_fp32_order2")
    assert r["kernel"] == "k_walk_step"
    assert set(per_kernel) == {"k_walk_step", "k_finalize+gaps"}
    B = bench.alg_bytes(4, rain_field=False)["k_step_fused"]
    assert r["alg_bytes_per_cell"] == B == 38
    assert r["achieved"] == pytest.approx(B * CELLS / 0.070e-3 / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0)
    # the committed summaries hold both figures for this kernel and workload
    assert r["traffic"] is not None and 0.5 * B * CELLS < r["traffic"] < 2.0 * B * CELLS
    assert r["valu_lane_instr_per_cell"] is not None
    vpeak, _ = bench.valu_peak()
    assert r["valu_issue_frac"] == pytest.approx(r["valu_lane_instr_per_cell"] * CELLS / 0.070e-3 / vpeak)
    assert r["limiter"] == ("valu" if r["valu_issue_frac"] > r["frac"] else "hbm")


def test_two_pass_double_picks_the_longer_pass(bench):
    per_kernel, r = bench.roofline_fields({"k_flux": 0.20, "k_update": 0.05}, 0.26, 8, False, CELLS,
                                          "rainfall_runoff_2048_fp64_order2")
    assert r["kernel"] == "k_flux"
    assert per_kernel == {"k_flux": 0.20, "k_update": 0.05}
    assert r["alg_bytes_per_cell"] == bench.alg_bytes(8, rain_field=False)["k_flux"]
    assert r["traffic"] is not None and r["valu_lane_instr_per_cell"] is not None


def test_unprofiled_workload_leaves_the_counter_figures_empty(bench):
    _, r = bench.roofline_fields({"k_step_fused": 0.5, "k_finalize+gaps": 0.01}, 0.52, 4, True, CELLS,
                                 "rainfall_runoff_333_fp32_order1")
    assert r["traffic"] is None and r["valu_lane_instr_per_cell"] is None and r["valu_issue_frac"] is None
    assert r["limiter"] == "hbm"
    assert r["alg_bytes_per_cell"] == bench.alg_bytes(4, rain_field=True)["k_step_fused"]


def test_kernel_longer_than_its_step_is_refused(bench):
    with pytest.raises(AssertionError):
        bench.roofline_fields({"k_step_fused": 0.09, "k_finalize+gaps": 0.0}, 0.08, 4, False, CELLS,
                              "rainfall_runoff_2048_fp32_order2")


def test_more_than_one_gpu_is_refused_before_any_gpu_work():
    """`bench.py --gpus 2` (what a multi-GPU launch would pass) exits with
    status 2 and one JSON line naming the reason, before importing torch."""
    import json
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_default_run_is_the_contract_one(bench, monkeypatch):
    """No flags: one GPU, 1000 timed steps after 50 warm-up steps, the 2048^2
    FP32 order-2 rainfall-runoff workload, the ensemble and adjoint objects
    included."""
    monkeypatch.setattr("sys.argv", ["bench.py"])
    a = bench.parse()
    assert (a.gpus, a.steps, a.warmup, a.size, a.case, a.dtype, a.order) == (1, 1000, 50, 2048, "runoff", "fp32", 2)
    assert a.ensemble_seconds > 0 and not a.no_adjoint
