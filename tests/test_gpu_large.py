"""The large-detector workload against the FP64 oracle: 1024^2 x 512 voxels at
0.4 mm from 12288 views of 96 x 896 (2304 views per turn, 16 blocks of 32
slices), on the GPU in full and on the oracle at sampled slices.

The oracle rebins and filters each sampled block's whole window of
n_angle_add = 3292 views (the reference's per-block loop) and backprojects the
sampled slices with all their rows: block 0 slices 3 and 28, block 15 slice 17.
These run the backprojector's wide-window instantiation (LDS pitch 105) and
its large tables, which only this shape reaches.

plain, fast_f32   <= 1e-4 relative RMSE
mixed_f16/bf16    <= 1 HU RMSE
"""
import numpy as np
import pytest
import torch

from listctrecons_amd import recon, scan
from oracle import oracle as O
from helpers import hu_errors, oracle_plan, rel_rmse

pytestmark = pytest.mark.gpu

WORKLOAD = "helical_96x896_1024x512"
SAMPLES = {0: [3, 28], 15: [17]}


@pytest.fixture(scope="module")
def large(ctx):
    w = scan.workload(WORKLOAD)
    print("large detector: synthesising", flush=True)
    fan, a, c = recon.synth_scan(ctx, w)
    pl = recon.plan_helical(w.prm(), a, c)
    assert pl.n_blocks == 16 and pl.n_slices_recon == 512
    yield w, fan, a, c, pl
    del fan
    torch.cuda.empty_cache()


@pytest.fixture(scope="module")
def large_ref(large):
    w, fan, a, c, _ = large
    opl, _, _ = oracle_plan(w)
    fan_np = fan.cpu().numpy()
    refs = {}
    for b, zs in SAMPLES.items():
        print(f"oracle: block {b} slices {zs}", flush=True)
        v = O.reconstruct(opl, fan_np, a, c, window=w.window, blocks=(b, b + 1), iz_list=zs)
        refs[b] = np.stack([v[32 * b + z] for z in zs])
    del fan_np
    return refs


def sampled(vol, refs):
    got = np.concatenate([np.stack([vol[32 * b + z] for z in SAMPLES[b]]) for b in refs])
    ref = np.concatenate([refs[b] for b in refs])
    return got, ref


@pytest.fixture(scope="module")
def plain_volume(ctx, large):
    w, fan, a, c, pl = large
    print("large detector: plain", flush=True)
    v = ctx.reconstruct(pl, recon.preset("plain", window=w.window), a, c, fan).clone()
    ctx.synchronize()
    return v


@pytest.mark.parametrize("preset", ["plain", "fast_f32", "fast_f32_x6", "mixed_f16", "mixed_bf16"])
def test_large_detector_sampled_parity(ctx, large, large_ref, plain_volume, preset):
    w, fan, a, c, pl = large
    print(f"large detector: {preset}", flush=True)
    if preset == "plain":
        vol = plain_volume
    else:
        vol = ctx.reconstruct(pl, recon.preset(preset, window=w.window), a, c, fan)
        ctx.synchronize()
    got, ref = sampled(vol.cpu().numpy(), large_ref)
    assert np.isfinite(ref).all()
    err = rel_rmse(got, ref)
    rmse_hu, max_hu = hu_errors(got, ref)
    print(f"  rel rmse {err:.2e}  rmse {rmse_hu:.4f} HU  max {max_hu:.3f} HU", flush=True)
    if preset in ("plain", "fast_f32", "fast_f32_x6"):
        assert err <= 1e-4
    else:
        assert rmse_hu <= 1.0


PI_SLICES = [3, 28]


@pytest.fixture(scope="module")
def large_ref_pi(large):
    w, fan, a, c, _ = large
    opl, _, _ = oracle_plan(w)
    fan_np = fan.cpu().numpy()
    print(f"oracle: PI-original block 0 slices {PI_SLICES}", flush=True)
    v = O.reconstruct(opl, fan_np, a, c, rebin_kind=2, weight_kind=1, bp_kind=1,
                      window=w.window, blocks=(0, 1), iz_list=PI_SLICES)
    del fan_np
    return np.stack([v[z] for z in PI_SLICES])


@pytest.mark.parametrize("preset", ["fast_f32_pi_ori", "mixed_f16_pi_ori"])
def test_large_detector_pi_original(ctx, large, large_ref_pi, preset):
    """The PI-original stage set with the folded backprojector at the
    wide-window instantiation (LDS pitch 105) against the oracle's
    PI-original reconstruction of block 0: FP32 <= 1e-4, f16 <= 1 HU."""
    w, fan, a, c, pl = large
    print(f"large detector: {preset}", flush=True)
    vol = ctx.reconstruct(pl, recon.preset(preset, window=w.window), a, c, fan)
    ctx.synchronize()
    got = np.stack([vol[z].cpu().numpy() for z in PI_SLICES])
    del vol
    ref = large_ref_pi
    assert np.isfinite(ref).all()
    err = rel_rmse(got, ref)
    rmse_hu, max_hu = hu_errors(got, ref)
    print(f"  rel rmse {err:.2e}  rmse {rmse_hu:.4f} HU  max {max_hu:.3f} HU", flush=True)
    if preset.startswith("fast_f32"):
        assert err <= 1e-4
    else:
        assert rmse_hu <= 1.0


def test_large_detector_paths_agree(ctx, large, plain_volume):
    """Over the whole 1024^2 x 512 volume: the pair-folded resident
    backprojector against the plain per-voxel path.  Different kernels, no
    shared backprojection code."""
    w, fan, a, c, pl = large
    ref = plain_volume
    for preset, tol in (("fast_f32", 1e-4), ("mixed_f16", None)):
        print(f"large detector, whole volume: {preset}", flush=True)
        got = ctx.reconstruct(pl, recon.preset(preset, window=w.window), a, c, fan)
        ctx.synchronize()
        d = (got - ref).double()
        if tol is not None:
            assert float(d.norm() / ref.double().norm()) <= tol
        else:
            assert float(d.pow(2).mean().sqrt()) * scan.hu_per_unit() <= 1.0
        del got, d


def test_large_detector_slab_equals_the_whole_volume(ctx, large):
    """A 4-block slab from a buffer holding only its views equals those
    blocks of the one-shot volume bit for bit (mixed f16: the block's pairs
    are split over four workgroups the same way in both)."""
    w, fan, a, c, pl = large
    opts = recon.preset("mixed_f16", window=w.window)
    whole = ctx.reconstruct(pl, opts, a, c, fan).clone()
    v0, v1 = recon.slab_views(pl, c, 4, 8)
    got = ctx.reconstruct_slabs(pl, opts, a, c, fan[v0:v1].contiguous(), v0, 4, 8)
    ctx.synchronize()
    assert torch.equal(got, whole[32 * 4:32 * 8])
