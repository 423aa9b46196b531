"""Benchmark: seconds per volume of the helical FDK path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--path NAME]

A step is one complete reconstruction (rebin -> filter -> backproject of all
planned blocks) of the 512 x 512 x 256 clinical workload from 8192 views of a
64 x 736 detector that are already resident in HBM.  The headline path is
the mixed-precision one (f16 filtered projections, FP32 accumulation); FP32
storage on the same resident pipeline, bf16 storage and the plain FP32 path
(the reference's per-block loop on the GPU) are timed beside it, and the FP64
oracle is timed on the host on a bounded sample.  The large-detector workload
(1024^2 x 512 from 12288 views of 96 x 896) is timed after it on the same
path, in one piece and streamed in z-slabs (--large, --large-slab).
Optional: slab streaming of the clinical workload from pinned host memory
(--stream-slabs), the learned correction's inference (--no-learned to skip),
the LISTCTRecons CLI end to end on the workload's files (--cli-e2e).

The roofline of the backprojector is taken against the unit that binds it,
the VALU: the busy cycles the committed counters of the launch give
(profiles/: SQ_ACTIVE_INST_VALU less the cycles two VALU instructions shared,
SQ_ACTIVE_INST_VALU2) over the SIMD-cycles of the kernel's launch time, which
this run measures with HIP events.

Prints one JSON line (rank 0).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "s/volume + backproj GUPS, 512²×256: FP32 vs mixed/learned at ≤1 HU RMSE"
HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E
ENGINE_HZ = 2.4e9  # MI355X peak engine clock
N_CU, N_SIMD = 256, 1024
# VALU issue: a wave64 VALU instruction holds its SIMD for one quad-cycle (two
# for transcendentals), and two waves' FP32 add / mul / fma (and 32-bit
# add / logic) instructions can share one (SQ_ACTIVE_INST_VALU2 counts those
# quad-cycles): 2 cycles per instruction at best, 4 for conversions, fract,
# floor, min / max / med3, shifts, packed and mixed 16-bit arithmetic, 8 for
# rcp / sqrt (listctrecons_amd/bin/issue_rates measures them).  The VALU is
# busy 4 (ACTIVE_INST_VALU - ACTIVE_INST_VALU2) cycles per launch out of
# 1024 SIMDs x 2.4 GHz x the launch time.  At 2 cycles per instruction the
# lane-operation peak would be 1024 x 64 / 2 x 2.4 GHz = 78.6e12 / s.
VALU_CYCLES_PEAK = N_SIMD * ENGINE_HZ
VALU_LANE_OPS_PEAK = N_SIMD * 64 / 2 * ENGINE_HZ
KERNEL_SOURCE = os.path.join(ROOT, "listctrecons_amd", "csrc", "lct_resident.hip")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_path(ctx, pl, opts, a, c, fan, vol, steps, warmup):
    import torch
    for _ in range(warmup):
        ctx.reconstruct(pl, opts, a, c, fan, vol)
    ctx.synchronize()
    # whole-job throughput: `steps` volumes back to back, one wait at the end
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.reconstruct(pl, opts, a, c, fan, vol)
    ctx.synchronize()
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    # stage breakdown: a second pass in profiling mode, which times each stage
    # with HIP events and waits for it
    ctx.set_profiling(True)
    stage = {"rebin_ms": 0.0, "filter_ms": 0.0, "backproject_ms": 0.0}
    bp_kernel_ms = 0.0
    launches = 0
    per = []
    for _ in range(steps):
        s0 = time.perf_counter()
        ctx.reconstruct(pl, opts, a, c, fan, vol)
        t = ctx.timings()
        per.append(time.perf_counter() - s0)
        for k in stage:
            stage[k] += t[k]
        bp_kernel_ms += t["backproject_kernel_ms"]
        launches += t["n_backproject_launches"]
    torch.cuda.synchronize()
    ctx.set_profiling(False)
    updates = t["backproject_updates"]
    return {
        "s_per_volume": total / steps,
        "median_s": sorted(per)[len(per) // 2],
        "stage_ms": {k: v / steps for k, v in stage.items()},
        "bp_ms_per_launch": stage["backproject_ms"] / max(launches, 1),
        # the backprojector kernel alone (without the pass adding split blocks)
        "bp_kernel_ms_per_launch": bp_kernel_ms / max(launches, 1),
        "bp_launches_per_volume": launches / steps,
        "updates": updates,
    }


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_pi_ori(w, fan_np, a, c, sample_slices):
    """The FP64 oracle's PI-original stage set on block 0 (rebin, weight and
    filter of its window, `sample_slices` backprojected) on all host threads:
    the volume (NaN outside the sample) and its time scaled to a whole volume."""
    from oracle import oracle as O
    p = O.Prm()
    pp = w.prm()
    for f, _ in O.Prm._fields_:
        setattr(p, f, getattr(pp, f))
    pl = O.plan_helical(p, a, c)
    kinds = dict(rebin_kind=2, weight_kind=1, bp_kind=1, window=w.window, blocks=(0, 1))
    t0 = time.perf_counter()
    O.reconstruct(pl, fan_np, a, c, iz_list=[], **kinds)
    t_rf = time.perf_counter() - t0
    t0 = time.perf_counter()
    vol = O.reconstruct(pl, fan_np, a, c, iz_list=list(sample_slices), **kinds)
    t_slice = max(time.perf_counter() - t0 - t_rf, 0.0) / len(sample_slices)
    per_volume = pl.n_blocks * (t_rf + 32 * t_slice)
    threads = O.num_threads()
    return ({"value": round(per_volume, 3), "unit": "s/volume", "cores": threads, "kind": "port",
             "sample": f"FP64 oracle, {threads} threads, block 0 of {pl.n_blocks}: rebin+weight+"
                       f"filter of its window ({t_rf:.2f} s) and {len(sample_slices)} of 32 slices "
                       f"backprojected ({t_slice:.2f} s/slice), scaled to {pl.n_blocks} blocks x 32 "
                       f"slices"}, vol)


def cpu_baseline(w, fan_np, a, c, sample_slices=tuple(range(0, 32, 2))):
    """FP64 oracle on the host: block 0 (rebin + filter of its whole window, and
    `sample_slices` backprojected), scaled to a whole volume, on all host
    threads; and a single-thread figure from smaller sub-samples."""
    from oracle import oracle as O
    p = O.Prm()
    pp = w.prm()
    for f, _ in O.Prm._fields_:
        setattr(p, f, getattr(pp, f))
    pl = O.plan_helical(p, a, c)
    threads = O.num_threads()
    t0 = time.perf_counter()
    O.reconstruct(pl, fan_np, a, c, window=w.window, blocks=(0, 1), iz_list=[])
    t_rf = time.perf_counter() - t0
    t0 = time.perf_counter()
    vol = O.reconstruct(pl, fan_np, a, c, window=w.window, blocks=(0, 1),
                        iz_list=list(sample_slices))
    t_all = time.perf_counter() - t0
    t_slice = max(t_all - t_rf, 0.0) / len(sample_slices)
    per_volume = pl.n_blocks * (t_rf + 32 * t_slice)

    # single thread: rebin of block 0's window, filter of 256 of its rows,
    # 64 rows of one slice backprojected; each scaled to the volume
    g, ig = pl.ctg, pl.ig
    s, _, _, _ = O.block_info(pl, c, 0)
    first = s + pl.data_offset
    win = fan_np[first:first + g.n_angle_add]
    taps = O.ramp_taps(g.ns, g.ds, w.window)
    O.set_num_threads(1)
    try:
        t0 = time.perf_counter()
        par = O.rebin(g, win, kind=1)
        t_rebin1 = time.perf_counter() - t0
        rows = par.reshape(-1, g.ns)
        t0 = time.perf_counter()
        O.filter_rows(taps, rows[:256])
        t_filter1 = (time.perf_counter() - t0) * rows.shape[0] / 256
        # the backprojection's cost does not depend on the values it sums
        betas = a[s + g.add_projections:s + g.add_projections + g.n_angle + g.n_proj_turn // 2 + 1]
        t0 = time.perf_counter()
        O.backproject(g, ig, par, betas, iz_list=[11], iy_range=(224, 288))
        t_slice1 = (time.perf_counter() - t0) * ig.ny / 64
    finally:
        O.set_num_threads(threads)
    per_volume_1 = pl.n_blocks * (t_rebin1 + t_filter1 + 32 * t_slice1)
    return {
        "value": per_volume,
        "unit": "s/volume",
        "cores": threads,
        "kind": "port",
        "cpu": cpu_model(),
        "nproc": os.cpu_count(),
        "sample": (f"FP64 oracle, {threads} threads, block 0 of {pl.n_blocks}: rebin+filter of its "
                   f"{pl.ctg.n_angle_add}-view window ({t_rf:.2f} s) and {len(sample_slices)} of "
                   f"32 slices backprojected ({t_slice:.2f} s/slice), scaled to "
                   f"{pl.n_blocks} blocks x 32 slices"),
        "sample_seconds": round(t_rf + t_all, 3),
        "single_thread": {
            "value": per_volume_1, "unit": "s/volume",
            "sample": (f"1 thread: rebin of block 0's window ({t_rebin1:.2f} s), filter of 256 "
                       f"rows scaled to {rows.shape[0]} ({t_filter1:.1f} s), 64 rows of slice 11 "
                       f"scaled to {ig.ny} rows ({t_slice1:.1f} s/slice); x 32 slices x "
                       f"{pl.n_blocks} blocks")},
    }, vol, sample_slices


def stream_slabs(ctx, pl, opts, a, c, host_fan, vol, blocks_per_slab, reps=2):
    """Reconstruct slab by slab from pinned host memory: each slab's views are
    staged on a copy stream (double buffered) while the previous slab is
    backprojected.  The views a slab shares with the previous one are copied
    from that slab's buffer on the device; only the new views cross PCIe.
    Returns seconds per volume (best of `reps`), the slab count, the largest
    slab's view count and the views uploaded per volume."""
    import torch
    from listctrecons_amd import recon
    g = pl.ctg
    view = g.nt * g.ns
    slabs = []
    for b0 in range(0, pl.n_blocks, blocks_per_slab):
        b1 = min(b0 + blocks_per_slab, pl.n_blocks)
        slabs.append((b0, b1) + recon.slab_views(pl, c, b0, b1))
    max_views = max(v1 - v0 for _, _, v0, v1 in slabs)
    bufs = [torch.empty(max_views * view, dtype=torch.float32, device="cuda") for _ in range(2)]
    # two non-default streams (default priority): reconstruction and copies
    comp = torch.cuda.Stream()
    copy = torch.cuda.Stream()
    torch.cuda.synchronize()
    ctx.set_stream(comp.cuda_stream)
    flat = host_fan.view(-1)
    best = None
    try:
        for _ in range(reps):
            copied = [torch.cuda.Event(), torch.cuda.Event()]
            freed = [torch.cuda.Event(), torch.cuda.Event()]
            for e in freed:
                e.record(comp)
            torch.cuda.synchronize()
            t0 = time.perf_counter()

            uploaded = [0]

            def issue(k):
                b0, b1, v0, v1 = slabs[k]
                with torch.cuda.stream(copy):
                    copy.wait_event(freed[k % 2])
                    dst = bufs[k % 2]
                    have = 0  # leading views of this slab already on the device
                    if k > 0:
                        _, _, pv0, pv1 = slabs[k - 1]
                        if pv0 <= v0 < pv1:
                            have = min(pv1, v1) - v0
                            src = bufs[(k - 1) % 2]
                            dst[:have * view].copy_(src[(v0 - pv0) * view:(v0 - pv0 + have) * view],
                                                    non_blocking=True)
                    dst[have * view:(v1 - v0) * view].copy_(flat[(v0 + have) * view:v1 * view],
                                                            non_blocking=True)
                    uploaded[0] += v1 - v0 - have
                    copied[k % 2].record(copy)

            issue(0)
            for k, (b0, b1, v0, v1) in enumerate(slabs):
                if k + 1 < len(slabs):
                    issue(k + 1)
                comp.wait_event(copied[k % 2])
                part = bufs[k % 2][:(v1 - v0) * view].view(v1 - v0, g.nt, g.ns)
                ctx.reconstruct_slabs(pl, opts, a, c, part, v0, b0, b1,
                                      vol[32 * b0:32 * b1])
                freed[k % 2].record(comp)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
    finally:
        ctx.set_stream(0)
    return best, len(slabs), max_views, uploaded[0]


def cli_end_to_end(w, host_fan, head_vol, path):
    """The drop-in CLI (FDKReconstructor through liblct) on the reference's files:
    writes proj.raw, angles.raw, couch.raw and scan.prm to a scratch directory,
    runs `LISTCTRecons scan.prm <factory string>` and times the process (reading
    the files, HIP initialisation, reconstruction, writing image.raw and
    rebin.raw).  Its image must equal the library's volume."""
    import shutil
    import subprocess
    import tempfile
    import numpy as np
    from listctrecons_amd import build, scan
    words = {"mixed_f16": "f16", "mixed_bf16": "bf16", "fast_f32": "resident"}
    if path not in words:
        raise SystemExit(f"--cli-e2e: no factory string for path {path}")
    window = {1: "hann-", 2: "shepp-"}.get(int(w.window), "")
    rtype = f"Helicalfdkramprebingpu-{window}{words[path]}"
    d = tempfile.mkdtemp(prefix="lct_cli_", dir=os.environ.get("TMPDIR", "/tmp"))
    try:
        prm = scan.write_scan_files(w, d, host_fan.numpy())
        t0 = time.perf_counter()
        r = subprocess.run([build.CLI, prm, rtype], stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True)
        dt = time.perf_counter() - t0
        if r.returncode != 0:
            raise SystemExit(f"CLI failed ({r.returncode}): {r.stdout[-2000:]}")
        img = np.fromfile(os.path.join(d, "image.raw"), np.float32).reshape(head_vol.shape)
        return {"s_per_volume": round(dt, 3), "factory": rtype,
                "identical_to_library": bool(np.array_equal(img, head_vol)),
                "note": "one LISTCTRecons process on files in a scratch directory: file "
                        "reads, HIP initialisation, reconstruction and image writes; the "
                        "files were just written, so they may be read from the page cache"}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def storage_of(path):
    """Filtered-projection storage of a preset: f16, bf16 or f32."""
    return {"mixed_f16": "f16", "mixed_bf16": "bf16", "resident_direct_f16": "f16",
            "mixed_f16_gather": "f16", "mixed_bf16_gather": "bf16"}.get(path, "f32")


def load_profile(kernel_tag):
    """Per-launch counter figures of a kernel on a workload, committed under
    profiles/ (rocprofv3 runs); {} when there are none."""
    path = os.path.join(ROOT, "profiles", f"{kernel_tag}.json")
    if not os.path.exists(path):
        return {}
    try:
        with open(path) as f:
            return json.load(f)
    except Exception:
        return {}


def source_sha16(path=KERNEL_SOURCE):
    import hashlib
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def backprojector_roofline(prof, kernel, updates, launches, launch_s, bytes_per_tap, n_vox,
                           rays=2, taps_from="LDS"):
    """Roofline of the backprojector launch against the unit that binds it,
    the VALU.  achieved = the VALU's busy cycles per launch, 4 x
    (SQ_ACTIVE_INST_VALU - SQ_ACTIVE_INST_VALU2) quad-cycles from the
    committed counters of this kernel and workload (they follow the code and
    the workload, not the timing), over the kernel's launch time measured in
    this run with HIP events on the context stream; peak = 1024 SIMDs x
    2.4 GHz of VALU cycles; frac = achieved / peak.  Beside it: the same
    instructions timed as if each took 2 cycles (`frac_at_2_cycles`, a lower
    bound), VALU lane-operations per update, LDS-active and bank-conflict
    fractions, measured HBM traffic (FETCH_SIZE / WRITE_SIZE, copy-calibrated)
    and its share of the 8 TB/s peak, and the gather-byte rate of the north
    star's definition (`rays` x 4 taps x element size per update; the folded
    kernel's taps come from LDS, not HBM).  The counters match the source when
    the committed hash is that of the file the profile names."""
    per_launch_updates = updates / launches
    gather_bytes = (updates * rays * 4 * bytes_per_tap + n_vox * 4) / launches
    out = {"bound": "valu", "kernel": kernel}
    insts = prof.get("sq_insts_valu_per_launch")
    act, act2 = prof.get("sq_active_inst_valu_per_launch"), prof.get("sq_active_inst_valu2_per_launch")
    cycles = launch_s * ENGINE_HZ
    if insts and act and act2 is not None:
        busy = 4.0 * (act - act2)
        out.update({
            "achieved": round(busy / launch_s / 1e12, 3),
            "peak": round(VALU_CYCLES_PEAK / 1e12, 3),
            "unit": "T VALU SIMD-cycles/s",
            "frac": round(busy / launch_s / VALU_CYCLES_PEAK, 4),
            "valu_busy_cycles_per_launch": busy,
            "valu_insts_per_launch": insts,
            "frac_at_2_cycles": round(insts * 64 / launch_s / VALU_LANE_OPS_PEAK, 4),
            "valu_lane_ops_per_update": round(insts * 64 / per_launch_updates, 2),
        })
    else:
        out.update({"achieved": None, "peak": round(VALU_CYCLES_PEAK / 1e12, 3),
                    "unit": "T VALU SIMD-cycles/s", "frac": None})
    traffic = prof.get("traffic_bytes_per_launch")
    out.update({
        "traffic": traffic,
        "traffic_frac_of_hbm_peak": round(traffic / launch_s / HBM_PEAK, 4) if traffic else None,
        "salu_insts_per_launch": prof.get("sq_insts_salu_per_launch"),
        "lds_busy": (round(prof["sq_lds_idx_active_per_launch"] / N_CU / cycles, 3)
                     if "sq_lds_idx_active_per_launch" in prof else None),
        "lds_bank_conflict": (round(prof["sq_lds_bank_conflict_per_launch"] / N_CU / cycles, 3)
                              if "sq_lds_bank_conflict_per_launch" in prof else None),
        "gather_GBps": round(gather_bytes / launch_s / 1e9, 1),
        "gather_definition": f"{rays} ray{'s' if rays > 1 else ''} x 4 bilinear taps x {bytes_per_tap} B "
                             f"per voxel-view update + 4 B per voxel written; "
                             f"{per_launch_updates:.4g} updates per launch; served from {taps_from}",
        "launch_ms": round(launch_s * 1e3, 3),
        "counters": "profiles/" + prof.get("sq_source", "") if prof else None,
        "counters_match_source": (prof.get("source_sha16") == source_sha16(
            os.path.join(ROOT, prof.get("source_file", os.path.relpath(KERNEL_SOURCE, ROOT)))))
        if prof else None,
    })
    return out


def large_workload(ctx, args, recon, scan, torch):
    """The large-detector workload (by default 1024^2 x 512 from 12288 views of
    96 x 896) on the headline path: 1 warm-up + 2 volumes back to back from
    HBM-resident projections, the stage breakdown and the backprojector's
    roofline, and z-slab streaming from pinned host memory."""
    wl = scan.workload(args.large)
    log(f"large workload {wl.name}: synthesising {wl.readings} views of {wl.nt}x{wl.ns}")
    fan, a, c = recon.synth_scan(ctx, wl)
    pl = recon.plan_helical(wl.prm(), a, c)
    vol = torch.empty((pl.n_slices_recon, pl.ig.ny, pl.ig.nx), dtype=torch.float32, device="cuda")
    opts = recon.preset(args.path, window=wl.window)
    r = time_path(ctx, pl, opts, a, c, fan, vol, 2, 1)
    log(f"  {r['s_per_volume']*1e3:.1f} ms/volume  stages {r['stage_ms']}")
    storage = "f32" if args.path.startswith("fast_f32") else ("bf16" if "bf16" in args.path else "f16")
    prof = load_profile(f"bp_fold_{storage}_{wl.name}")
    n_vox = pl.n_slices_recon * pl.ig.nx * pl.ig.ny
    res = {"workload": wl.name, "views": wl.readings, "detector": f"{wl.nt}x{wl.ns}",
           "volume": f"{pl.ig.nx}x{pl.ig.ny}x{pl.n_slices_recon}", "path": args.path,
           "steps": 2, "warmup": 1,
           "s_per_volume": round(r["s_per_volume"], 6),
           "stage_ms": {k: round(v, 3) for k, v in r["stage_ms"].items()},
           "backproject_gups": round(r["updates"] / (r["stage_ms"]["backproject_ms"] * 1e-3) / 1e9, 2),
           "roofline": backprojector_roofline(prof, prof.get("kernel", "bp_fold_kernel"),
                                              r["updates"], r["bp_launches_per_volume"],
                                              r["bp_kernel_ms_per_launch"] * 1e-3,
                                              4 if storage == "f32" else 2, n_vox)}
    if not args.no_cpu_baseline:
        res["cpu_baseline"], res["hu_vs_fp64_oracle"] = large_cpu_baseline(wl, fan, a, c, vol)
    if args.large_slab > 0:
        host_fan = torch.empty(fan.shape, dtype=fan.dtype, pin_memory=True)
        host_fan.copy_(fan)
        svol = torch.empty_like(vol)
        log(f"  streaming in slabs of {args.large_slab} blocks")
        st_s, n_slabs, max_views, n_up = stream_slabs(ctx, pl, opts, a, c, host_fan, svol,
                                                      args.large_slab, reps=1)
        res["slab_streaming"] = {"blocks_per_slab": args.large_slab, "slabs": n_slabs,
                                 "s_per_volume": round(st_s, 6), "max_views_per_slab": max_views,
                                 "views_uploaded": n_up,
                                 "identical_to_one_shot": bool(torch.equal(svol, vol))}
        log(f"  {st_s*1e3:.1f} ms/volume streamed")
        del host_fan, svol
    del fan, vol
    torch.cuda.empty_cache()
    return res


def large_cpu_baseline(wl, fan, a, c, vol, sample_slices=(5, 22)):
    """FP64 oracle on all host threads for the large workload, extrapolated:
    block 0's window rebinned and filtered and `sample_slices` of its 32
    slices backprojected, scaled to every block x 32 slices; and the GPU
    volume's error against the oracle on that sample."""
    import numpy as np
    from listctrecons_amd import scan
    from oracle import oracle as O
    p = O.Prm()
    pp = wl.prm()
    for f, _ in O.Prm._fields_:
        setattr(p, f, getattr(pp, f))
    pl = O.plan_helical(p, a, c)
    s, _, _, _ = O.block_info(pl, c, 0)
    n_views = s + pl.data_offset + pl.ctg.n_angle_add  # the views block 0 reads
    fan_np = fan[:n_views].cpu().numpy()
    threads = O.num_threads()
    log(f"  FP64 oracle on the host: block 0, slices {list(sample_slices)}")
    t0 = time.perf_counter()
    O.reconstruct(pl, fan_np, a, c, window=wl.window, blocks=(0, 1), iz_list=[])
    t_rf = time.perf_counter() - t0
    t0 = time.perf_counter()
    ovol = O.reconstruct(pl, fan_np, a, c, window=wl.window, blocks=(0, 1),
                         iz_list=list(sample_slices))
    t_all = time.perf_counter() - t0
    t_slice = max(t_all - t_rf, 0.0) / len(sample_slices)
    idx = list(sample_slices)
    ref = ovol[idx]
    got = vol[idx].cpu().numpy()
    err = hu_stats(got, ref, np, scan)
    err["rel_rmse"] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    err["sample"] = f"block 0 slices {idx}, all voxels"
    cb = {"value": round(pl.n_blocks * (t_rf + 32 * t_slice), 3), "unit": "s/volume",
          "cores": threads, "kind": "port", "extrapolated": True, "cpu": cpu_model(),
          "sample": (f"FP64 oracle, {threads} threads, block 0 of {pl.n_blocks}: rebin+filter of "
                     f"its {pl.ctg.n_angle_add}-view window ({t_rf:.2f} s) and {len(idx)} of 32 "
                     f"slices backprojected ({t_slice:.2f} s/slice), extrapolated to "
                     f"{pl.n_blocks} blocks x 32 slices"),
          "sample_seconds": round(t_rf + t_all, 3)}
    return cb, err


def hu_stats(x, ref, np, scan):
    d = (x - ref).astype(np.float64) * scan.hu_per_unit()
    return {"rmse": round(float(np.sqrt(np.mean(d * d))), 4),
            "max_abs": round(float(np.max(np.abs(d))), 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="helical_64x736_512x256")
    ap.add_argument("--path", default="mixed_f16")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip timing the learned correction's inference")
    ap.add_argument("--no-pi-ori", action="store_true",
                    help="skip the PI-original stage set (per-block FP32)")
    ap.add_argument("--variants", default="fast_f32,fast_f32_x6,mixed_bf16",
                    help="comma-separated presets to time besides the headline path "
                         "(e.g. fast_f32,fast_f32_gather,mixed_f16_gather,mixed_bf16)")
    ap.add_argument("--stream-slabs", type=int, default=2,
                    help="also time slab streaming from pinned host memory, N blocks per slab "
                         "(0: off)")
    ap.add_argument("--cli-e2e", action="store_true",
                    help="also time the LISTCTRecons CLI on the workload's files (file I/O, "
                         "process start and HIP initialisation included)")
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="run one device copy of this many GiB first (PMC scale reference)")
    ap.add_argument("--large", default="helical_96x896_1024x512",
                    help="also time this larger workload (1 warm-up + 2 volumes, and slab "
                         "streaming of --large-slab blocks); empty: off")
    ap.add_argument("--large-slab", type=int, default=4)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "runs on one GPU: the helical FDK path would shard by "
                                   "32-slice blocks (each GPU a block range and the views it "
                                   "reads, no collectives), but only the single-GPU path is "
                                   "built and measured; run with --gpus 1"}))
        sys.exit(2)

    import numpy as np
    import torch

    from listctrecons_amd import recon, scan

    w = scan.workload(args.workload)
    if args.copy_probe_gib > 0:
        # a plain device-to-device copy of known size: the HBM counters of this
        # dispatch calibrate the FETCH_SIZE / WRITE_SIZE readings of the others
        n = int(args.copy_probe_gib * (1 << 30)) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        torch.cuda.empty_cache()
    ctx = recon.Context(0)
    log(f"synthesising {w.name}: {w.readings} views of {w.nt}x{w.ns}")
    fan, a, c = recon.synth_scan(ctx, w)
    pl = recon.plan_helical(w.prm(), a, c)
    vol = torch.empty((pl.n_slices_recon, pl.ig.ny, pl.ig.nx), dtype=torch.float32,
                      device="cuda")

    head_opts = recon.preset(args.path, window=w.window)
    log(f"timing {args.path}: {args.warmup} warm-up + {args.steps} steps")
    head = time_path(ctx, pl, head_opts, a, c, fan, vol, args.steps, args.warmup)
    head_vol = vol.cpu().numpy()
    log(f"  {head['s_per_volume']*1e3:.2f} ms/volume  stages {head['stage_ms']}")

    # host <-> device transfers a caller holding the scan in host memory pays on
    # top of `value`: the projection file up (pinned), the volume down
    host_fan = torch.empty(fan.shape, dtype=fan.dtype, pin_memory=True)
    host_fan.copy_(fan)
    dev_fan = torch.empty_like(fan)
    host_vol = torch.empty(vol.shape, dtype=vol.dtype, pin_memory=True)
    h2d, d2h = [], []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dev_fan.copy_(host_fan, non_blocking=True)
        torch.cuda.synchronize()
        h2d.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        host_vol.copy_(vol, non_blocking=True)
        torch.cuda.synchronize()
        d2h.append(time.perf_counter() - t0)
    h2d_s, d2h_s = min(h2d), min(d2h)
    del dev_fan, host_vol

    streamed = None
    if args.stream_slabs > 0:
        log(f"timing slab streaming, {args.stream_slabs} blocks per slab")
        svol = torch.empty_like(vol)
        st_s, n_slabs, max_views, n_up = stream_slabs(ctx, pl, head_opts, a, c, host_fan, svol,
                                                      args.stream_slabs)
        streamed = {"blocks_per_slab": args.stream_slabs, "slabs": n_slabs,
                    "s_per_volume": round(st_s, 6),
                    "max_views_per_slab": max_views,
                    "views_uploaded": n_up,
                    "identical_to_one_shot": bool(torch.equal(svol, vol)),
                    "note": "projections streamed from pinned host memory: on a copy stream, "
                            "each slab's views shared with the previous slab are copied on the "
                            "device and the rest uploaded, while the previous slab is "
                            "reconstructed"}
        log(f"  {st_s*1e3:.1f} ms/volume streamed in {n_slabs} slabs")
        del svol

    cli = None
    if args.cli_e2e:
        cli = cli_end_to_end(w, host_fan, head_vol, args.path)
        log(f"  CLI end to end {cli['s_per_volume']:.3f} s/volume")
    del host_fan

    plain = None
    plain_vol = None
    if not args.no_plain:
        vol2 = torch.empty_like(vol)
        log("timing plain FP32 (per-block reference structure)")
        plain = time_path(ctx, pl, recon.preset("plain", window=w.window), a, c, fan, vol2,
                          max(1, min(args.steps, 3)), 1)
        plain_vol = vol2.cpu().numpy()
        log(f"  {plain['s_per_volume']*1e3:.2f} ms/volume  stages {plain['stage_ms']}")
        del vol2

    variants, variant_vols = {}, {}
    for name in [v for v in args.variants.split(",") if v and v != args.path]:
        log(f"timing variant {name}")
        vv = torch.empty_like(vol)
        r = time_path(ctx, pl, recon.preset(name, window=w.window), a, c, fan, vv,
                      max(1, min(args.steps, 3)), 1)
        variant_vols[name] = vv.cpu().numpy()
        variants[name] = {"s_per_volume": round(r["s_per_volume"], 6),
                          "stage_ms": {k: round(x, 3) for k, x in r["stage_ms"].items()},
                          "backproject_gups": round(r["updates"] / (r["stage_ms"]["backproject_ms"]
                                                                    * 1e-3) / 1e9, 2),
                          "hu_vs_headline": hu_stats(variant_vols[name], head_vol, np, scan)}
        if plain_vol is not None:
            variants[name]["hu_vs_fp32_plain"] = hu_stats(variant_vols[name], plain_vol, np, scan)
        del vv

    pi_ori, pi_vol, pi_plain_vol = None, None, None
    if not args.no_pi_ori:
        # the reference's PI-original stage set: its rebin (Rebiner.cu:156-233),
        # pre-weighting (PreWeighting.cu:8-44) and backprojector
        # (Backprojector.cu:171-307), on the headline path (view pairs folded:
        # view j where its ray lies in the voxel's PI window, else view j + h)
        # and on the reference's per-block FP32 loop
        pi_ori = {}
        runs = [("plain", "plain_pi_ori", "bp_pi_ori")]
        if args.path != "plain" and "gather" not in args.path:  # no folded gather form
            runs.insert(0, (args.path, f"{args.path}_pi_ori", "bp_fold_pi"))
        for key, preset, tag in runs:
            log(f"timing the PI-original stage set ({preset})")
            vp = torch.empty_like(vol)
            r = time_path(ctx, pl, recon.preset(preset, window=w.window), a, c, fan, vp,
                          max(1, min(args.steps, 3)), 1)
            if key == "plain":
                pi_plain_vol = vp.cpu().numpy()
            else:
                pi_vol = vp.cpu().numpy()
            del vp
            st = "f32" if key in ("plain", "fast_f32") else storage_of(key)
            prof_pi = load_profile(f"{tag}_{st}_{w.name}" if key != "plain" else f"{tag}_{w.name}")
            pi_ori[key] = {
                "s_per_volume": round(r["s_per_volume"], 6),
                "stage_ms": {k: round(x, 3) for k, x in r["stage_ms"].items()},
                "backproject_gups": round(r["updates"] / (r["stage_ms"]["backproject_ms"] * 1e-3)
                                          / 1e9, 2),
                "updates_per_volume": r["updates"],
                "roofline": backprojector_roofline(
                    prof_pi, prof_pi.get("kernel", "bp_pi_ori_literal_kernel" if key == "plain"
                                         else "bp_fold_kernel"), r["updates"],
                    r["bp_launches_per_volume"], r["bp_kernel_ms_per_launch"] * 1e-3,
                    4 if st == "f32" else 2, pl.n_slices_recon * pl.ig.nx * pl.ig.ny, rays=1,
                    taps_from="L2 / HBM" if key == "plain" else "LDS")}
            log(f"  {r['s_per_volume']*1e3:.2f} ms/volume  stages {r['stage_ms']}")
        if len(runs) == 2:
            pi_ori["speedup_vs_plain"] = round(pi_ori["plain"]["s_per_volume"]
                                               / pi_ori[args.path]["s_per_volume"], 3)
        pi_ori["note"] = ("updates = voxels x half a turn of views; the plain loop launches once "
                          "per 32-slice block, the folded backprojector once per volume")

    learned = None
    if not args.no_learned:
        # inference cost of the learned correction on this volume (the trained
        # network's accuracy is measured by python -m listctrecons_amd.learned)
        from listctrecons_amd import learned as L
        log("timing the learned correction's inference")
        net = recon.Cnn(ctx, seed=0)
        L.correct(net, vol)
        torch.cuda.synchronize()
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            L.correct(net, vol)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        net.close()
        learned = {"cnn_s_per_volume": round(min(ts), 6),
                   "s_per_volume_with_correction": round(head["s_per_volume"] + min(ts), 6),
                   "network": "3D residual CNN, 6 x (3x3x3 conv, 32 channels), BF16 MFMA",
                   "note": "inference time with initial weights (the cost does not depend on "
                           "them); trained accuracy in profiles/learned_pitch15.json"}
        torch.cuda.empty_cache()

    storage = storage_of(args.path)
    bytes_per_tap = 2 if storage in ("f16", "bf16") else 4
    updates = head["updates"]
    bp_s = head["bp_kernel_ms_per_launch"] * 1e-3
    launches = head["bp_launches_per_volume"]
    gups = updates / (head["stage_ms"]["backproject_ms"] * 1e-3) / 1e9
    prof = load_profile(f"bp_fold_{storage}_{w.name}")
    roof = backprojector_roofline(prof, prof.get("kernel", "bp_fold_kernel"), updates, launches,
                                  bp_s, bytes_per_tap, pl.n_slices_recon * pl.ig.nx * pl.ig.ny)

    out = {
        "metric": METRIC,
        "value": round(head["s_per_volume"], 6),
        "unit": "s/volume",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["s_per_volume"] * 1e3, 3),
        "median_s_per_volume_stage_timed": round(head["median_s"], 6),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "storage": storage,
        "data": "synthetic: analytic ellipsoid phantom line integrals + Poisson noise",
        "config": {"workload": w.name, "views": w.readings, "detector": f"{w.nt}x{w.ns}",
                   "volume": f"{pl.ig.nx}x{pl.ig.ny}x{pl.n_slices_recon}", "path": args.path},
        "backproject_gups": round(gups, 2),
        "stage_ms": {k: round(v, 3) for k, v in head["stage_ms"].items()},
        "roofline": roof,
        "pcie": {"h2d_ms": round(h2d_s * 1e3, 3),
                 "h2d_GBps": round(fan.numel() * 4 / h2d_s / 1e9, 1),
                 "d2h_ms": round(d2h_s * 1e3, 3),
                 "s_per_volume_incl_transfers": round(head["s_per_volume"] + h2d_s + d2h_s, 6),
                 "note": "projection file up from pinned host memory, volume down to pinned "
                         "host memory; not in value"},
    }
    if streamed:
        out["slab_streaming"] = streamed
    if cli:
        out["cli_end_to_end"] = cli
    if variants:
        out["variants"] = variants
    if pi_ori:
        out["pi_original"] = pi_ori
    if learned:
        out["learned"] = learned
    if plain:
        out["fp32_plain_s_per_volume"] = round(plain["s_per_volume"], 6)
        out["fp32_plain_stage_ms"] = {k: round(v, 3) for k, v in plain["stage_ms"].items()}
        out["speedup_vs_fp32_plain"] = round(plain["s_per_volume"] / head["s_per_volume"], 3)
        out["hu_vs_fp32_plain"] = hu_stats(head_vol, plain_vol, np, scan)
    if "fast_f32" in variants and args.path != "fast_f32":
        # the same resident pipeline and backprojector with FP32 storage and the
        # FP32 direct filter: what the reduced precision alone buys
        out["speedup_vs_fp32_resident"] = round(variants["fast_f32"]["s_per_volume"]
                                                / head["s_per_volume"], 3)

    if args.large and args.large != w.name:
        out["large_workload"] = large_workload(ctx, args, recon, scan, torch)

    if not args.no_cpu_baseline:
        log("timing the FP64 oracle on the host (bounded sample)")
        fan_np = fan.cpu().numpy()
        cb, ovol, sl = cpu_baseline(w, fan_np, a, c)
        out["cpu_baseline"] = cb
        idx = list(sl)
        ref = ovol[idx]
        vols = [("mixed", head_vol), ("fp32_plain", plain_vol)] + \
            [(k, v) for k, v in variant_vols.items()]
        for name, v in vols:
            if v is None:
                continue
            s = hu_stats(v[idx], ref, np, scan)
            s["rel_rmse"] = float(np.linalg.norm(v[idx] - ref) / np.linalg.norm(ref))
            s["sample"] = f"block 0 slices {idx}, all voxels"
            out[f"hu_vs_fp64_oracle_{name}"] = s
        if pi_vol is not None:
            log("the FP64 oracle's PI-original stage set on the same sample")
            ocb, opi = cpu_pi_ori(w, fan_np, a, c, idx)
            for key, v in ((args.path, pi_vol), ("plain", pi_plain_vol)):
                if v is None or key not in out["pi_original"]:
                    continue
                s = hu_stats(v[idx], opi[idx], np, scan)
                s["rel_rmse"] = float(np.linalg.norm(v[idx] - opi[idx]) / np.linalg.norm(opi[idx]))
                s["sample"] = f"block 0 slices {idx}, all voxels"
                out["pi_original"][key]["hu_vs_fp64_oracle"] = s
            out["pi_original"]["cpu_baseline"] = ocb
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
