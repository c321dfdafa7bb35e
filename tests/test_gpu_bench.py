"""bench.py's output contract: one JSON line with the fields the benchmark
report is read from, the roofline object (achieved = algorithmic bytes per
launch / launch time, frac = achieved / peak) and the CPU baseline beside
it.  Runs the headline workload at a small photon count in a child process."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_line_contract():
    args = ["--photons", "2e6", "--steps", "2", "--warmup", "1", "--cpu-seconds", "1", "--batch-scale", "0.002",
            "--serve-requests", "12", "--serve-pool-scale", "0.05"]
    print("bench.py " + " ".join(args), flush=True)
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                         timeout=240, cwd=ROOT)
    assert out.returncode == 0, out.stderr[-2000:]
    lines = [ln for ln in out.stdout.splitlines() if ln.startswith("{")]
    assert len(lines) == 1, out.stdout
    d = json.loads(lines[0])
    for k, t in (("metric", str), ("value", float), ("unit", str), ("n_gpus", int), ("steps", int),
                 ("warmup", int), ("ms_per_step", float), ("higher_is_better", bool), ("scaling", str),
                 ("dtype", str), ("data", str), ("config", dict), ("roofline", dict), ("cpu_baseline", dict)):
        assert isinstance(d[k], t), k
    assert "vs_baseline" in d and d["n_gpus"] == 1 and d["steps"] == 2 and d["warmup"] == 1
    assert d["unit"] == "photons/s" and d["higher_is_better"] is True and d["dtype"] == "f32"
    assert "workload" in d["config"] and d["config"]["photons_per_step"] == 2000000
    # value = photons per step / wall time per step
    assert d["value"] == pytest.approx(2e6 / (d["ms_per_step"] / 1e3), rel=1e-3)
    r = d["roofline"]
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert r["achieved"] == pytest.approx(r["bytes_per_step"] * r["steps_per_launch"] / (r["kernel_ms_per_launch"] / 1e3)
                                          / 1e9, rel=2e-3)
    assert r["frac"] == pytest.approx(r["achieved"] / r["peak"], rel=2e-3)
    assert r["traffic"] is None or r["traffic"] > 0
    v = d["vector_roofline"]
    assert v["unit"] == "TFLOP/s" and v["peak"] == 157.3
    assert v["frac"] == pytest.approx(v["achieved"] / v["peak"], rel=1e-2)
    c = d["cpu_baseline"]
    assert c["value"] > 0 and c["unit"] == "photons/s" and c["cores"] >= 1
    assert c["kind"] in ("reference", "port") and c["sample"]
    # the FP32 line keeps within the metric's 1 % of FP64 on the same streams
    assert d["fp32_vs_fp64_rel_l2"] < 1e-2
    # the metric's batched part: 256 instances in one replayed launch
    bt = d["batch"]
    assert bt["instances"] == 256 and bt["photons_per_s"] > 0 and bt["photons"] > 0
    assert bt["photons_per_s"] == pytest.approx(bt["photons"] / (bt["ms"] / 1e3), rel=1e-3)
    assert max(bt["fp32_vs_fp64_rel_l2_first4"]) < 1e-2
    # and its serving part: latencies and alone == in the mix
    sv = d["serve"]
    assert sv["requests"] == 12 and sv["photons_per_s"] > 0
    assert 0 < sv["latency_ms_p50"] <= sv["latency_ms_p99"] <= sv["latency_ms_max"]
    assert sv["alone_vs_mix_rel_l2_max"] is not None and sv["alone_vs_mix_rel_l2_max"] < 1e-9

