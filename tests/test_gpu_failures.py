"""Failure handling through the C ABI on the GPU: a device allocation that
cannot succeed returns LCT_ERR_NOMEM and leaves the context usable, and
non-finite projections stay confined to the blocks whose windows read them
(the reference checks neither).  Every kernel here exits normally."""
import ctypes as C

import numpy as np
import pytest
import torch

from listctrecons_amd import _lib, recon, scan
from oracle import oracle as O
from helpers import cpu_fan, oracle_plan

pytestmark = pytest.mark.gpu


def test_allocation_failure_is_reported_and_recoverable(ctx):
    p = C.c_void_p()
    rc = _lib.lib().lct_device_malloc(ctx.h, C.byref(p), C.c_size_t(1 << 50))  # 1 PiB
    assert rc == _lib.LCT_ERR_NOMEM and not p.value
    assert b"hipMalloc" in _lib.lib().lct_last_error()
    # the context still works
    w = scan.workload("tiny")
    a, c = w.angles_couch()
    pl = recon.plan_helical(w.prm(), a, c)
    vol = ctx.reconstruct(pl, recon.preset("fast_f32"), a, c, torch.from_numpy(cpu_fan(w)).cuda())
    ctx.synchronize()
    assert torch.isfinite(vol).all()


@pytest.fixture(scope="module")
def poisoned():
    """A 3-block scan with a NaN and an Inf in views that only block 2's
    voxels backproject (block 1's window ends at file view 917), and which
    blocks the FP64 oracle finds non-finite."""
    w = scan.workload("tiny", readings=1400, pitch=1.7)
    pl, a, c = oracle_plan(w)
    fan = cpu_fan(w).copy()
    fan[940, 3, 20] = np.nan
    fan[950, 5, 40] = np.inf
    ref = O.reconstruct(pl, fan, a, c, window=w.window)
    finite = [bool(np.isfinite(ref[32 * b:32 * b + 32]).all()) for b in range(pl.n_blocks)]
    assert finite == [True, True, False]
    return w, fan, finite


@pytest.mark.parametrize("preset", ["plain", "fast_f32", "mixed_f16"])
def test_non_finite_views_stay_in_their_blocks(ctx, poisoned, preset):
    w, fan, finite = poisoned
    a, c = w.angles_couch()
    pl = recon.plan_helical(w.prm(), a, c)
    print(f"non-finite input, {preset}", flush=True)
    vol = ctx.reconstruct(pl, recon.preset(preset, window=w.window), a, c,
                          torch.from_numpy(fan).cuda())
    ctx.synchronize()
    v = vol.cpu().numpy()
    assert [bool(np.isfinite(v[32 * b:32 * b + 32]).all()) for b in range(3)] == finite


def test_unsupported_shapes_are_refused_before_any_launch(ctx):
    """Shapes a kernel cannot take are refused on the host with
    LCT_ERR_UNSUPPORTED, and the context keeps working."""
    # reduced-precision storage needs the pair-folded backprojector (even npt)
    w = scan.workload("tiny", n_proj_turn=97, readings=1200)
    a, c = w.angles_couch()
    pl = recon.plan_helical(w.prm(), a, c)
    fan = torch.from_numpy(cpu_fan(w)).cuda()
    print("odd npt with f16 storage", flush=True)
    with pytest.raises(_lib.LctError) as e:
        ctx.reconstruct(pl, recon.preset("mixed_f16", window=w.window), a, c, fan)
    assert e.value.code == _lib.LCT_ERR_UNSUPPORTED and "pair-folded" in str(e.value)
    # the Toeplitz filter needs ns % 4 == 0
    ns, nt = 30, 8
    rows = torch.zeros(nt * ns, device="cuda")
    q = torch.zeros((ns + 3) * (nt + 8), dtype=torch.float16, device="cuda")
    with pytest.raises(_lib.LctError) as e:
        ctx.filter_to_bp_layout(_lib.FILTER_TOEPLITZ_BF16X3, _lib.F16, ns, nt,
                                np.zeros(2 * ns, np.float32), rows, 1, q)
    assert e.value.code == _lib.LCT_ERR_UNSUPPORTED
    # the direct FP32 filter holds one row per workgroup: ns <= 2048
    g = _lib.CTGeom()
    g.ns, g.nt, g.n_angle, g.ds = 2050, 1, 1, 1.0
    with pytest.raises(_lib.LctError) as e:
        ctx.filter(g, np.zeros(2 * g.ns, np.float32), torch.zeros(g.ns, device="cuda"), 1)
    assert e.value.code == _lib.LCT_ERR_UNSUPPORTED
    vol = ctx.reconstruct(pl, recon.preset("fast_f32", window=w.window), a, c, fan)
    ctx.synchronize()
    assert torch.isfinite(vol).all()
