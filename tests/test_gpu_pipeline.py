"""Whole-pipeline parity of lct_reconstruct with the FP64 oracle.

plain      the reference's per-block loop on the GPU, FP32: <= 1e-4 relative RMSE
fast_f32   HBM-resident views, pair-folded backprojector, FP32: <= 1e-4
mixed_*    f16 / bf16 filtered projections, FP32 accumulation: <= 1 HU RMSE
At the 512^2 x 256 clinical size the oracle runs on sampled slices and rows;
size-independent properties (linearity, water scale) cover the full volume.
"""
import numpy as np
import pytest
import torch

from listctrecons_amd import _lib, recon, scan
from oracle import oracle as O
from helpers import cpu_fan, hu_errors, oracle_plan, rel_rmse

pytestmark = pytest.mark.gpu


def gpu_recon(ctx, w, fan_np_or_dev, preset, plan=None, **kw):
    a, c = w.angles_couch()
    pl = plan or recon.plan_helical(w.prm(), a, c)
    fan = fan_np_or_dev
    if isinstance(fan, np.ndarray):
        fan = torch.from_numpy(fan).cuda()
    opts = recon.preset(preset, window=w.window)
    for k, v in kw.items():
        setattr(opts, k, v)
    vol = ctx.reconstruct(pl, opts, a, c, fan)
    ctx.synchronize()
    return vol.cpu().numpy()


@pytest.fixture(scope="module")
def cfg1():
    w = scan.workload("helical_16x368_256x32")
    pl, a, c = oracle_plan(w)
    fan = cpu_fan(w)
    ref = O.reconstruct(pl, fan, a, c, window=w.window)
    return w, fan, ref


@pytest.mark.parametrize("preset,tol", [("plain", 1e-4), ("fast_f32", 1e-4),
                                        ("fast_f32_gather", 1e-4), ("fast_f32_toeplitz", 2e-4),
                                        ("fast_f32_x6", 1e-4)])
def test_fp32_paths_match_oracle(ctx, cfg1, preset, tol):
    w, fan, ref = cfg1
    print(f"{preset} on {w.name}", flush=True)
    got = gpu_recon(ctx, w, fan, preset)
    err = rel_rmse(got, ref)
    rmse_hu, max_hu = hu_errors(got, ref)
    print(f"  rel rmse {err:.2e}  rmse {rmse_hu:.4f} HU  max {max_hu:.3f} HU", flush=True)
    assert err <= tol


@pytest.mark.parametrize("preset", ["resident_direct_f16", "mixed_f16", "mixed_bf16",
                                    "mixed_f16_gather"])
def test_reduced_storage_within_1hu(ctx, cfg1, preset):
    w, fan, ref = cfg1
    print(f"{preset} on {w.name}", flush=True)
    got = gpu_recon(ctx, w, fan, preset)
    rmse_hu, max_hu = hu_errors(got, ref)
    print(f"  rmse {rmse_hu:.4f} HU  max {max_hu:.3f} HU  rel {rel_rmse(got, ref):.2e}",
          flush=True)
    assert rmse_hu <= 1.0


@pytest.fixture(scope="module")
def cfg1_pi():
    w = scan.workload("helical_16x368_256x32")
    pl, a, c = oracle_plan(w)
    fan = cpu_fan(w)
    ref = O.reconstruct(pl, fan, a, c, rebin_kind=2, weight_kind=1, bp_kind=1, window=w.window)
    return w, fan, ref


@pytest.mark.parametrize("preset,hu", [("plain_pi_ori", None), ("fast_f32_pi_ori", None),
                                       ("mixed_f16_pi_ori", 1.0), ("mixed_bf16_pi_ori", 1.0)])
def test_pi_original_paths_match_oracle(ctx, cfg1_pi, preset, hu):
    """The PI-original stage set against the oracle's: the reference's
    per-block loop and the folded resident backprojector (view j of a pair
    where its ray lies in the voxel's PI window, else view j + h) within
    1e-4 relative RMSE in FP32, within 1 HU with 16-bit storage."""
    w, fan, ref = cfg1_pi
    print(f"{preset} on {w.name}", flush=True)
    got = gpu_recon(ctx, w, fan, preset)
    err = rel_rmse(got, ref)
    rmse_hu, max_hu = hu_errors(got, ref)
    print(f"  rel rmse {err:.2e}  rmse {rmse_hu:.4f} HU  max {max_hu:.3f} HU", flush=True)
    if hu is None:
        assert err <= 1e-4
    else:
        assert rmse_hu <= hu


VARIANTS = [
    ("tiny", {}, {}),
    ("tiny", {"z_dir": -1}, {}),
    ("tiny", {"view_offset": 3, "data_offset": 5, "offset_x": 1.5, "offset_y": -2.0}, {}),
    ("tiny", {"pitch": 1.7, "readings": 1400}, {}),
    ("tiny", {}, {"rebin_kind": 0}),
    ("tiny", {}, {"rebin_kind": 2, "weight_kind": 1, "bp_kind": 1}),
    ("tiny", {"z_dir": -1}, {"rebin_kind": 2, "weight_kind": 1, "bp_kind": 1}),
    ("tiny", {"pitch": 1.7, "readings": 1400}, {"rebin_kind": 2, "weight_kind": 1, "bp_kind": 1}),
    ("tiny", {"dx": 14.0}, {"rebin_kind": 2, "weight_kind": 1, "bp_kind": 1}),
    ("tiny", {"n_proj_turn": 97, "readings": 1200}, {}),  # odd npt: no pair folding
    # volume corners at |p_hat| / dso ~ 0.6 (the backprojector's asin
    # polynomial) and ~ 0.73 (beyond its range: the library asin)
    ("tiny", {"dx": 14.0}, {}),
    ("tiny", {"dx": 17.0}, {}),
]


@pytest.mark.parametrize("preset", ["plain", "fast_f32"])
@pytest.mark.parametrize("name,wkw,okw", VARIANTS)
def test_variants(ctx, preset, name, wkw, okw):
    w = scan.workload(name, **wkw)
    print(f"{preset} {name} {wkw} {okw}", flush=True)
    pl, a, c = oracle_plan(w)
    fan = cpu_fan(w)
    ref = O.reconstruct(pl, fan, a, c, rebin_kind=okw.get("rebin_kind", 1),
                        weight_kind=okw.get("weight_kind", 0), window=w.window,
                        bp_kind=okw.get("bp_kind", 0))
    got = gpu_recon(ctx, w, fan, preset, **okw)
    assert rel_rmse(got, ref) <= 1e-4


@pytest.mark.parametrize("name,wkw,preset", [
    ("tiny", {"readings": 1400, "pitch": 1.7}, "plain"),
    ("tiny", {"readings": 1400, "pitch": 1.7}, "fast_f32"),
    ("tiny", {"n_proj_turn": 97, "readings": 1200}, "fast_f32"),  # resident, not folded
])
def test_last_block_filtered_views(ctx, name, wkw, preset):
    """lct_last_block_filtered == the oracle's rebinned + filtered window of
    the last block (the reference's rebin.raw)."""
    w = scan.workload(name, **wkw)
    pl, a, c = oracle_plan(w)
    fan = cpu_fan(w)
    gpu_recon(ctx, w, fan, preset)
    got = ctx.last_block_filtered(recon.plan_helical(w.prm(), a, c)).cpu().numpy()
    g = pl.ctg
    s, _, _, _ = O.block_info(pl, c, pl.n_blocks - 1)
    first = s + pl.data_offset
    q = O.filter_rows(O.ramp_taps(g.ns, g.ds, w.window), O.rebin(g, fan[first:first + g.n_angle_add]))
    assert rel_rmse(got, q) < 2e-6


def test_missing_views_raise(ctx):
    w = scan.workload("tiny")
    a, c = w.angles_couch()
    pl = recon.plan_helical(w.prm(), a, c)
    short = torch.zeros((w.readings // 2, w.nt, w.ns), device="cuda")
    with pytest.raises(_lib.LctError):
        ctx.reconstruct(pl, recon.preset("plain"), a, c, short)


# ------------------------------------------------------------ clinical size

@pytest.fixture(scope="module")
def clinical(ctx):
    w = scan.workload("helical_64x736_512x256")
    fan, a, c = recon.synth_scan(ctx, w)
    return w, fan, a, c


SAMPLE_BLOCKS = (0, 5)
SAMPLE_SLICES = [3, 30]
SAMPLE_ROWS = (200, 264)


@pytest.fixture(scope="module")
def clinical_ref(clinical):
    w, fan, a, c = clinical
    pl, _, _ = oracle_plan(w)
    fan_np = fan.cpu().numpy()
    refs = {}
    for b in SAMPLE_BLOCKS:
        print(f"oracle block {b}", flush=True)
        v = O.reconstruct(pl, fan_np, a, c, window=w.window, blocks=(b, b + 1),
                          iz_list=SAMPLE_SLICES, iy_range=SAMPLE_ROWS)
        refs[b] = v
    return refs


def sampled(vol, refs):
    g, r = [], []
    for b, v in refs.items():
        for z in SAMPLE_SLICES:
            sl = slice(*SAMPLE_ROWS)
            g.append(vol[32 * b + z, sl])
            r.append(v[32 * b + z, sl])
    return np.stack(g), np.stack(r)


@pytest.mark.parametrize("preset", ["plain", "fast_f32", "resident_direct_f16", "mixed_f16",
                                    "mixed_bf16"])
def test_clinical_sampled_parity(ctx, clinical, clinical_ref, preset):
    w, fan, a, c = clinical
    print(f"{preset} on {w.name}", flush=True)
    vol = gpu_recon(ctx, w, fan, preset)
    got, ref = sampled(vol, clinical_ref)
    err = rel_rmse(got, ref)
    rmse_hu, max_hu = hu_errors(got, ref)
    print(f"  rel rmse {err:.2e}  rmse {rmse_hu:.4f} HU  max {max_hu:.3f} HU", flush=True)
    if preset in ("plain", "fast_f32"):
        assert err <= 1e-4
    else:
        assert rmse_hu <= 1.0


@pytest.mark.parametrize("preset", ["plain", "fast_f32", "mixed_f16"])
def test_slabs_equal_the_whole_volume(ctx, clinical, preset):
    """lct_reconstruct_slabs from a buffer holding only the slab's views gives
    the same slices, bit for bit, as the whole-volume reconstruction."""
    w, fan, a, c = clinical
    print(f"slabs {preset}", flush=True)
    pl = recon.plan_helical(w.prm(), a, c)
    opts = recon.preset(preset, window=w.window)
    whole = ctx.reconstruct(pl, opts, a, c, fan).clone()
    for b0, b1 in ((0, 1), (2, 5), (7, 8)):
        v0, v1 = recon.slab_views(pl, c, b0, b1)
        part = fan[v0:v1].contiguous()
        got = ctx.reconstruct_slabs(pl, opts, a, c, part, v0, b0, b1)
        ctx.synchronize()
        assert torch.equal(got, whole[32 * b0:32 * b1]), (b0, b1)
    # a buffer that lacks the slab's first view is refused
    v0, v1 = recon.slab_views(pl, c, 2, 5)
    with pytest.raises(_lib.LctError) as e:
        ctx.reconstruct_slabs(pl, opts, a, c, fan[v0 + 1:v1].contiguous(), v0 + 1, 2, 5)
    assert e.value.code == _lib.LCT_ERR_RANGE


@pytest.mark.parametrize("preset", ["fast_f32", "mixed_f16", "mixed_f16_gather"])
def test_back_to_back_volumes_equal_one_at_a_time(ctx, clinical, preset):
    """Host tables reach the device through the context's pinned staging ring
    without waits, so volumes can be enqueued back to back.  Three volumes
    enqueued before one wait, with the context on a non-default stream, equal
    bit for bit the volume of a profiled run (which waits after every
    stage)."""
    w, fan, a, c = clinical
    print(f"back to back {preset}", flush=True)
    pl = recon.plan_helical(w.prm(), a, c)
    assert pl.n_blocks == 8
    opts = recon.preset(preset, window=w.window)
    ctx.set_profiling(True)
    try:
        serial = ctx.reconstruct(pl, opts, a, c, fan).clone()
    finally:
        ctx.set_profiling(False)
    got = ctx.reconstruct(pl, opts, a, c, fan)
    ctx.synchronize()
    assert torch.equal(got, serial)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    ctx.set_stream(s.cuda_stream)
    try:
        outs = [torch.empty_like(serial) for _ in range(3)]
        for o in outs:
            ctx.reconstruct(pl, opts, a, c, fan, o)
        ctx.synchronize()
    finally:
        ctx.set_stream(0)
    for o in outs:
        assert torch.equal(o, serial)


def test_clinical_linearity_full_volume(ctx, clinical):
    """reconstruct(2 p) == 2 reconstruct(p) over the whole 512^2 x 256 volume."""
    w, fan, a, c = clinical
    v1 = gpu_recon(ctx, w, fan, "fast_f32")
    v2 = gpu_recon(ctx, w, fan * 2.0, "fast_f32")
    assert rel_rmse(v2, 2.0 * v1) < 1e-6
    assert np.isfinite(v1).all()
