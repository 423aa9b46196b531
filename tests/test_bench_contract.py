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
                  "--no-learned", "--variants", "", "--stream-slabs", "0", "--large", "",
                  timeout=300)
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
    # the VALU binds the backprojector; this workload has no committed
    # counters, so achieved / frac are null and the measured rates remain
    assert rf["bound"] == "valu" and rf["unit"] == "T VALU SIMD-cycles/s"
    assert abs(rf["peak"] - 1024 * 2.4e9 / 1e12) < 1e-3
    assert rf["launch_ms"] > 0 and rf["gather_GBps"] > 0
    cb = d["cpu_baseline"]
    assert cb["kind"] == "port" and cb["value"] > 0 and cb["cores"] >= 1 and cb["sample"]
    # the mixed path against the FP64 oracle and the plain path against both
    assert d["hu_vs_fp64_oracle_mixed"]["rmse"] <= 1.0
    assert d["hu_vs_fp64_oracle_fp32_plain"]["rel_rmse"] <= 1e-4
    assert d["speedup_vs_fp32_plain"] > 1.0
    # the PI-original stage set: timed, its roofline's measured rates, and
    # against the oracle's PI-original reconstruction of the same sample
    pi = d["pi_original"]
    for key in ("mixed_f16", "plain"):
        assert pi[key]["s_per_volume"] > 0 and pi[key]["backproject_gups"] > 0
        assert pi[key]["roofline"]["bound"] == "valu" and pi[key]["roofline"]["launch_ms"] > 0
    assert "L2" in pi["plain"]["roofline"]["gather_definition"]
    assert pi["plain"]["hu_vs_fp64_oracle"]["rel_rmse"] <= 1e-4
    assert pi["mixed_f16"]["hu_vs_fp64_oracle"]["rmse"] <= 1.0
    assert pi["speedup_vs_plain"] > 1.0
    assert pi["cpu_baseline"]["value"] > 0 and pi["cpu_baseline"]["kind"] == "port"


def test_roofline_recomputes_from_the_counters():
    """frac = 4 (SQ_ACTIVE_INST_VALU - SQ_ACTIVE_INST_VALU2) VALU-busy cycles
    / 1024 SIMDs / (launch s x 2.4 GHz), never below the 2-cycles-per-
    instruction figure; the LDS and HBM fractions from the same committed
    counters, for both committed workloads' profiles."""
    sys.path.insert(0, ROOT)
    import bench
    for name in ("helical_64x736_512x256", "helical_96x896_1024x512"):
        prof = bench.load_profile(f"bp_fold_f16_{name}")
        assert prof, name
        assert prof["source_sha16"] == bench.source_sha16(), \
            "backprojector source changed since its counters were committed"
        launch_s = 0.030
        rf = bench.backprojector_roofline(prof, prof["kernel"], 7.731e10, 1, launch_s, 2,
                                          512 * 512 * 256)
        busy = 4 * (prof["sq_active_inst_valu_per_launch"] - prof["sq_active_inst_valu2_per_launch"])
        frac = busy / 1024 / (launch_s * 2.4e9)
        assert rf["frac"] == pytest.approx(frac, rel=1e-3)
        assert rf["achieved"] / rf["peak"] == pytest.approx(frac, rel=1e-2)
        frac2 = prof["sq_insts_valu_per_launch"] * 2 / 1024 / (launch_s * 2.4e9)
        assert rf["frac_at_2_cycles"] == pytest.approx(frac2, rel=1e-3)
        # two instructions share a quad-cycle at most: busy >= 2 cycles each
        assert busy >= 2 * prof["sq_insts_valu_per_launch"] * (1 - 1e-6)
        assert rf["lds_busy"] == pytest.approx(
            prof["sq_lds_idx_active_per_launch"] / 256 / (launch_s * 2.4e9), abs=1e-3)
        assert rf["traffic_frac_of_hbm_peak"] == pytest.approx(
            prof["traffic_bytes_per_launch"] / launch_s / 8e12, abs=1e-4)


def test_large_cpu_baseline_is_labelled_extrapolated():
    """The large workload's CPU figure: block 0's window rebinned and filtered
    and two sampled slices backprojected by the FP64 oracle, extrapolated to
    the whole volume and labelled so; the volume's error on that sample (here
    the oracle's own slices in float32: rounding only)."""
    import numpy as np
    import torch
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import bench
    from helpers import cpu_fan, oracle_plan
    from listctrecons_amd import scan
    from oracle import oracle as O
    w = scan.workload("helical_16x368_256x32")
    pl, a, c = oracle_plan(w)
    fan = cpu_fan(w)
    ovol = O.reconstruct(pl, fan, a, c, window=w.window, blocks=(0, 1), iz_list=[5, 22])
    vol = torch.from_numpy(np.nan_to_num(ovol).astype(np.float32))
    cb, err = bench.large_cpu_baseline(w, torch.from_numpy(fan), a, c, vol)
    assert cb["extrapolated"] is True and cb["kind"] == "port" and cb["value"] > 0
    assert cb["cores"] >= 1 and "extrapolated to 1 blocks x 32 slices" in cb["sample"]
    assert err["rmse"] < 1e-3 and err["rel_rmse"] < 1e-6
    assert err["sample"] == "block 0 slices [5, 22], all voxels"
