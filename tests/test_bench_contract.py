"""bench.py's output contract: one JSON line with the metric, the roofline of
the backprojector and the CPU baseline; more than one GPU is refused."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args, timeout=600):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=timeout, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = run_bench("--gpus", "2", timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


@pytest.mark.gpu
def test_bench_line_fields():
    """The 256^2 x 32 workload (one block) with the plain path and the FP64
    oracle beside it: every field the contract names is present and sane."""
    print("bench on helical_16x368_256x32", flush=True)
    r = run_bench("--workload", "helical_16x368_256x32", "--steps", "2", "--warmup", "1",
                  "--no-learned", "--variants", "", "--stream-slabs", "0", timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["unit"] == "s/volume" and d["higher_is_better"] is False
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["warmup"] == 1
    assert d["value"] > 0 and abs(d["ms_per_step"] - d["value"] * 1e3) < 1e-2
    assert d["config"]["workload"] == "helical_16x368_256x32"
    rf = d["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert rf["achieved"] > 0 and abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-3
    cb = d["cpu_baseline"]
    assert cb["kind"] == "port" and cb["value"] > 0 and cb["cores"] >= 1 and cb["sample"]
    # the mixed path against the FP64 oracle and the plain path against both
    assert d["hu_vs_fp64_oracle_mixed"]["rmse"] <= 1.0
    assert d["hu_vs_fp64_oracle_fp32_plain"]["rel_rmse"] <= 1e-4
    assert d["speedup_vs_fp32_plain"] > 1.0
