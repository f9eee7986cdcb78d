"""bench.py's JSON line on the device: a short run of the default workload (the D3Q27 FP64
Y-bifurcation in a 512x256x256 box) in this process, checked field by field against the line's
contract (one GPU, whole-job MLUPS, the HBM roofline of the step kernel, the CPU oracle's rate) and
against its own arithmetic."""
import json
import sys

import pytest

import bench

pytestmark = pytest.mark.gpu

STEPS, WARMUP = 40, 4


def test_bench_line(monkeypatch, capsys):
    monkeypatch.setattr(sys, "argv", ["bench.py", "--steps", str(STEPS), "--warmup", str(WARMUP), "--learned", "0",
                                      "--ref-cadence-steps", "20", "--cpu-budget", "2",
                                      "--copy-bytes", str(64 << 20)])
    bench.main()
    lines = [ln for ln in capsys.readouterr().out.splitlines() if ln.startswith("{")]
    assert len(lines) == 1
    d = json.loads(lines[0])

    assert d["n_gpus"] == 1 and d["steps"] == STEPS and d["warmup"] == WARMUP
    assert d["higher_is_better"] is True and d["scaling"] == "weak" and d["vs_baseline"] is None
    assert d["unit"] == "MLUPS" and d["dtype"] == "f64" and d["metric"]
    c = d["config"]
    assert c["workload"] == "ybif-D3Q27-fp64-512x256x256"
    assert c["fluid_nodes"] == 968846  # the generated Y-bifurcation's lattice nodes
    assert c["streaming"] == "in place, one lattice" and c["population_stores"] == "plain"
    # value: whole-job fluid-node updates per second over the timed steps
    n = c["fluid_nodes"]
    assert d["value"] > 0
    assert d["value"] == pytest.approx(n / (d["ms_per_step"] * 1e-3) / 1e6, rel=1e-3)

    r = d["roofline"]
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert r["bytes_per_flup"] == 2 * 27 * 8
    assert r["achieved"] == pytest.approx(n * r["bytes_per_flup"] / (r["launch_avg_us"] * 1e-6) / 1e9, rel=1e-3)
    assert r["frac"] == pytest.approx(r["achieved"] / r["peak"], abs=1e-4)
    assert 0.1 < r["frac"] < 1.0
    # measured HBM bytes per launch from the committed counter summary: at least the algorithmic bytes
    assert r["traffic"] >= n * r["bytes_per_flup"]
    assert r["copy_kernel_GBs"] > 0

    for cpu, cores in ((d["cpu_baseline"], 1), (d["cpu_baseline_threads"], None)):
        assert cpu["unit"] == "MLUPS" and cpu["kind"] == "port" and cpu["value"] > 0 and cpu["sample"]
        assert cpu["cores"] == cores if cores else cpu["cores"] >= 1
    ref = d["reference_cadence"]
    assert ref["wss_every"] == 1 and ref["steps"] == 20 and ref["value"] > 0
