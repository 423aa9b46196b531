"""lct_backproject_noalloc, the allocation-free variant of lct_backproject
(Backprojector::backproject, CUDA/Backprojector.cu:101-169 and :252-307),
against lct_backproject through the C ABI: the same kernels on the same
tables, so the volumes are equal bit for bit.  Two calls in a row share the
library's fixed per-GPU table array, the second with a smaller block (fewer
angles, smaller volume), and the second result must not carry anything of
the first.  A block whose angle table exceeds the fixed array is refused
before any launch."""
import numpy as np
import pytest
import torch

from listctrecons_amd import _lib, recon, scan

pytestmark = pytest.mark.gpu


def block_inputs(name, seed):
    w = scan.workload(name)
    a, c = w.angles_couch()
    pl = recon.plan_helical(w.prm(), a, c)
    s = recon.block_info(pl, c, 0)[0]
    g, ig = pl.ctg, pl.ig
    betas = recon.block_betas(pl, a, s)
    rng = np.random.default_rng(seed)
    q = torch.from_numpy(rng.normal(size=(g.n_angle, g.nt, g.ns)).astype(np.float32)).cuda()
    gg = _lib.CTGeom.from_buffer_copy(bytes(g))
    igg = _lib.ImgGeom.from_buffer_copy(bytes(ig))
    igg.nz = 32
    return gg, igg, q, betas


@pytest.mark.parametrize("kind", [0, 1])
def test_noalloc_backprojector_equals_the_allocating_one(ctx, kind):
    # the larger block first, then a smaller one on the same table array
    for name, seed in (("helical_16x368_256x32", 1), ("tiny", 2)):
        g, ig, q, betas = block_inputs(name, seed)
        print(f"noalloc backprojector kind={kind} on {name}: n_angle {g.n_angle}, "
              f"{ig.nx}x{ig.ny}x32", flush=True)
        ref = torch.empty((32, ig.ny, ig.nx), dtype=torch.float32, device="cuda")
        got = torch.full_like(ref, float("nan"))
        ctx.backproject(kind, g, ig, q, betas, ref)
        ctx.backproject_noalloc(kind, g, ig, q, betas, got)
        ctx.synchronize()
        assert torch.isfinite(ref).all()
        assert torch.equal(got, ref), float((got - ref).abs().max())


def test_noalloc_backprojector_refuses_tables_beyond_its_array(ctx):
    g, ig, _, _ = block_inputs("tiny", 3)
    g.n_angle = 16384
    g.n_angle_add = 16384 + 2 * g.add_projections
    nb = g.n_angle + g.n_proj_turn // 2 + 1
    q = torch.zeros((g.n_angle, g.nt, g.ns), dtype=torch.float32, device="cuda")
    img = torch.zeros((32, ig.ny, ig.nx), dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.LctError) as e:
        ctx.backproject_noalloc(0, g, ig, q, np.zeros(nb, np.float32), img)
    assert e.value.code == _lib.LCT_ERR_UNSUPPORTED
    assert "16384" in str(e.value)
    ctx.synchronize()
    assert not img.any()
