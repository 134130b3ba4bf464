"""bench.py's output contract on the GPU: one JSON line with the metric,
throughput, step counts, the dominant kernel's roofline and the CPU baseline
(a small 2D shock/inclusion level, run as a child process)."""
from __future__ import annotations

import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_json_contract():
    print("\n[case] bench.py --nx 256 --ny 128 --steps 3 --warmup 1", flush=True)
    cmd = [sys.executable, os.path.join(ROOT, "bench.py"), "--nx", "256", "--ny", "128", "--box", "64",
           "--steps", "3", "--warmup", "1", "--cpu-seconds", "1"]
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["metric"] == "Mcell-updates/s on 1 GPU" and d["unit"] == "Mcell-updates/s"
    assert d["n_gpus"] == 1 and d["steps"] == 3 and d["warmup"] == 1 and d["higher_is_better"] is True
    assert d["value"] > 0 and d["ms_per_step"] > 0 and d["status"] == 0
    assert abs(d["value"] - 256 * 128 / (d["ms_per_step"] * 1e3)) <= 0.02 * d["value"]
    assert d["dtype"] == "f64" and "workload" in d["config"]
    rf = d["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in rf, k
    assert rf["bound"] == "hbm" and rf["peak"] == 8000.0 and rf["unit"] == "GB/s"
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-3
    cb = d["cpu_baseline"]
    assert cb["value"] > 0 and cb["kind"] == "port" and cb["cores"] >= 1 and cb["sample"]
    m = d["mixed"]
    assert m["value"] > 0 and m["l1_rel_err_vs_fp64"]["rho"] <= 1e-5


def test_bench_companion_runs():
    """--companions on: the 3D plate impact's mixed/FP64 figures and the 2D
    plate impact's gradient cost arrive in the same JSON line, from child runs
    at small sizes here (32^3, 64^2)."""
    print("\n[case] bench.py --companions on --n3 32 --n2g 64", flush=True)
    cmd = [sys.executable, os.path.join(ROOT, "bench.py"), "--nx", "256", "--ny", "128", "--box", "64",
           "--steps", "3", "--warmup", "1", "--no-cpu", "--companions", "on", "--n3", "32", "--n2g", "64"]
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    assert len(lines) == 1
    c = json.loads(lines[0])["companions"]
    p3, g = c["plate3d"], c["plate2d_grad"]
    assert "error" not in p3, p3
    assert "error" not in g, g
    assert p3["fp64"] > 0 and p3["mixed"] > 0 and p3["status"] == [0, 0] and p3["steps"] == 40
    assert abs(p3["mixed_over_fp64"] - p3["mixed"] / p3["fp64"]) <= 0.01 * p3["mixed_over_fp64"]
    assert p3["l1_rel_err_vs_fp64"]["rho"] <= 1e-5
    assert g["directions"] == 6 and g["forward_ms"] > 0 and g["gradient_ms"] > g["forward_ms"]
    assert abs(g["gradient_over_forward"] - g["gradient_ms"] / g["forward_ms"]) <= 0.01 * g["gradient_over_forward"]
    assert g["status"] == [0, 0]


def test_machine_peaks_are_physical():
    """The streaming and FMA kernels of libtenjin_peaks.so report rates that a
    device can deliver: positive, and at most the MI355X datasheet figures
    (8 TB/s HBM3E; 78.6 TFLOP/s FP64 vector; FP32 without packed
    instructions at most the packed 157.3) over a 1 GiB buffer, which no
    cache holds."""
    sys.path.insert(0, ROOT)
    import bench

    print("\n[case] machine peaks, 1 GiB buffers", flush=True)
    p = bench.machine_peaks(1 << 30)
    assert "error" not in p, p
    for k in ("hbm_read_gbs", "hbm_write_gbs", "hbm_copy_gbs"):
        assert isinstance(p[k], float) and 500.0 < p[k] <= 8000.0, (k, p[k])
    assert isinstance(p["fp64_fma_tflops"], float) and 5.0 < p["fp64_fma_tflops"] <= 78.6
    assert isinstance(p["fp32_fma_tflops"], float) and 5.0 < p["fp32_fma_tflops"] <= 157.3
