"""The benchmark's command-line contract: one JSON line of the agreed keys,
one GPU only."""

import json
import os
import subprocess
import sys

import pytest

from conftest import ROOT

BENCH = os.path.join(ROOT, "bench.py")


def test_bench_runs_on_one_gpu_only():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and not r.stdout.strip()


@pytest.mark.gpu
def test_bench_prints_one_json_line_of_the_contract():
    """A short run of the benchmark workload (600 x 200 x 50, f32, with the
    16-bit lines, no CPU sample) prints exactly one JSON line with the
    agreed keys, a healthy state and a roofline record of the dominant
    kernel."""
    r = subprocess.run([sys.executable, BENCH, "--steps", "6", "--warmup", "2",
                        "--cpu-seconds", "0", "--device-warmup", "0.05"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    assert len(lines) == 1, r.stdout
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["unit"] == "MLUPS" and d["n_gpus"] == 1 and d["steps"] == 6 and d["warmup"] == 2
    assert d["value"] > 0 and d["higher_is_better"] is True and d["dtype"] == "f32"
    assert d["config"]["grid"] == [600, 200, 50] and d["config"]["markers"] == 8200
    rl = d["roofline"]
    assert rl["bound"] == "hbm" and rl["unit"] == "GB/s" and rl["peak"] == 8000.0
    assert 0 < rl["frac"] < 1.0 and abs(rl["frac"] - rl["achieved"] / rl["peak"]) < 1e-3
    # the split step in one launch: every cell, the region after the markers' forces
    assert rl["cells_per_launch"] == 600 * 200 * 50 and rl["kernel"].startswith("k_step_fused")
    assert rl["bytes_per_launch"] == pytest.approx(152 * rl["cells_per_launch"], rel=1e-6)
    assert d["health"]["finite"] and d["health"]["rho_min"] > 0.9
    assert set(d["storage_16bit"]) == {"f16c", "f16", "bf16"}
    assert all(v["finite"] and v["value"] > 0 for v in d["storage_16bit"].values())
