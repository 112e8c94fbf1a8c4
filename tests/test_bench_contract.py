"""bench.py's output contract: one JSON line on stdout with the metric, the
roofline of the dominant kernel and the CPU baseline fields; one GPU only; the
algorithmic byte counts and the committed traffic figures it reports."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_algorithmic_bytes_per_cell():
    # uniform-mesh path: J and the metric terms are not loaded
    assert bench.kernel_alg_bytes("x_sweep", True, True) == 80.0      # Q 40 + RHS write 40
    assert bench.kernel_alg_bytes("y_sweep", True, True) == 120.0     # Q 40 + RHS read and write 80
    # z sweep: Q 40, RHS read 40, Q_out 40, Q^n at two of three stages, viscous fluxes 96
    assert bench.kernel_alg_bytes("z_sweep", True, True) == pytest.approx(40 + 40 + 40 + 80 / 3 + 96)
    assert bench.kernel_alg_bytes("z_sweep", False, True) == pytest.approx(40 + 40 + 40 + 80 / 3)
    assert bench.kernel_alg_bytes("viscous_flux", True, True) == 136.0  # Q 40 + fv write 96
    # metric fields: J 8 + three metric terms 24 per sweep, all ten for the viscous flux,
    # the xi and eta metrics once in three stages for the z sweep's eigenvalues
    assert bench.kernel_alg_bytes("x_sweep", True) == 112.0
    assert bench.kernel_alg_bytes("y_sweep", True) == 152.0
    assert bench.kernel_alg_bytes("z_sweep", True) == pytest.approx(40 + 40 + 40 + 80 / 3 + 96 + 32 + 16)
    assert bench.kernel_alg_bytes("viscous_flux", True) == 216.0
    # a fused step: stage 0 reads Q and writes Q1, stages 1 and 2 also read Q^n;
    # each stage reads J and the nine metric terms once (none on the uniform path)
    assert bench.ALG_BYTES_PER_STEP == (40 + 80 + 40) + 2 * (40 + 80 + 40 + 40)
    assert bench.ALG_BYTES_PER_STEP_UNIFORM == (40 + 40) + 2 * (40 + 40 + 40)


@pytest.mark.parametrize("variant", ["exact", "fast"])
@pytest.mark.parametrize("cls", ["x_sweep", "y_sweep", "z_sweep", "viscous_flux"])
def test_committed_traffic_matches_its_summary(cls, variant):
    cells = 256 ** 3
    t = bench.read_traffic(cls, 256, variant)
    assert t is not None
    path = os.path.join(ROOT, "profiles", "%s_iso256_%s_traffic.txt" % (bench.KERNEL_OF_CLASS[cls], variant))
    fetch = write = None
    for line in open(path):
        if line.startswith("FETCH_SIZE avg:"):
            fetch = float(line.split()[-2])
        if line.startswith("WRITE_SIZE avg:"):
            write = float(line.split()[-2])
    # the per-launch bytes are the per-cell fetch and write figures over the 256^3 cells
    assert t / cells == pytest.approx(fetch + write, rel=2e-3)
    # measured traffic is at least the compulsory bytes and within 2x of them
    alg = bench.kernel_alg_bytes(cls, True, True)
    assert 0.95 * alg <= t / cells <= 2.0 * alg
    assert bench.read_traffic(cls, 32, variant) is None


@pytest.mark.parametrize("variant", ["exact", "fast"])
def test_committed_valu_summary(variant):
    """The FP64 VALU figures the bench line carries come from the committed
    counter summary: every plain stage kernel is listed, its FP64 rate is its
    FP64 instructions (FMA as two flops) over its duration, and the step's
    FP64 instructions per cell-update are three stages of the four kernels."""
    ks = ["k_sweep_%s<0, 0, true, 0, 32, false, false>", "k_sweep_%s<1, 0, true, 0, 32, false, false>",
          "k_sweep_%s<2, 0, true, 1, 16, false, false>", "k_visc_flux_%s<3, true>"]
    vs = [bench.read_valu(k % variant, 256, variant) for k in ks]
    assert all(v is not None for v in vs)
    for v in vs:
        assert 0.0 < v["kernel_fp64_frac"] < 1.0 and 0.0 < v["kernel_valu_issue_frac"] <= 1.0
        assert v["kernel_fp64_tflops"] == pytest.approx(v["kernel_fp64_frac"] * 78.6, rel=0.01)
    step = 3 * sum(v["kernel_fp64_lane_ops_per_cell"] for v in vs)
    assert vs[0]["step_fp64_lane_ops_per_cell_update"] == pytest.approx(step, abs=4)
    assert bench.read_valu(ks[0] % variant, 32, variant) is None


@pytest.mark.gpu
def test_bench_json_line():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--n", "32", "--steps", "2", "--warmup", "1",
                        "--develop", "0", "--no-cpu-baseline"],
                       capture_output=True, text=True, timeout=600, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    assert d["metric"] == bench.METRIC and d["unit"] == "cell-updates/s"
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["warmup"] == 1
    assert d["value"] > 0 and d["ms_per_step"] > 0
    assert d["higher_is_better"] is True and d["scaling"] == "weak" and d["vs_baseline"] is None
    assert d["dtype"] == "f64" and d["data"] == "synthetic"
    assert "workload" in d["config"]
    rf = d["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == bench.HBM_PEAK_GBS
    assert rf["achieved"] > 0 and rf["frac"] == pytest.approx(rf["achieved"] / rf["peak"])
    assert rf["traffic"] is None  # no committed counter summary for 32^3
    assert d["cpu_baseline"] is None
    # the roofline fraction follows the algorithmic bytes of a stage launch:
    # a third of the step's bytes per cell-update for every cell
    cells = d["config"]["cells"]
    assert rf["alg_bytes_per_launch"] == pytest.approx(rf["step_alg_bytes_per_cell_update"] / 3.0 * cells)
    assert rf["achieved"] == pytest.approx(rf["alg_bytes_per_launch"] / (rf["avg_launch_ms"] * 1e-3) / 1e9)
    assert rf["step_frac"] == pytest.approx(rf["step_alg_bytes_per_cell_update"] * d["plain"]["value"] / 1e9
                                            / rf["peak"])
    # value is the fastest method's among those within the accuracy bar (the
    # density fluctuation's rms) against plain WENO5-JS after the same steps
    p = d["plain"]
    best = p["value"]
    for key in ("decisions_every_step", "decisions_every_other_step"):
        h = d["hybrid"][key]
        assert h["l2_over_fluctuation_rms"] >= h["l2_over_density_rms"] >= 0.0
        if h["l2_over_fluctuation_rms"] <= 1e-3:
            best = max(best, h["value"])
    assert d["value"] == pytest.approx(best)
    assert d["config"]["value_method"].startswith("plain" if best == p["value"] else "hybrid")
    # the other kernel variant is timed for the record only, and the sweep
    # tile widths are the autotuner's choice for each direction
    o = d["plain_other_variant"]
    assert o["variant"] == ("fast" if d["config"]["variant"] == "exact" else "exact")
    assert o["value"] > 0 and o["ms_per_step"] > 0
    assert len(p["sweep_faces_per_tile_line"]) == 3
    assert all(nf in (16, 32) for nf in p["sweep_faces_per_tile_line"])
