"""Stage-by-stage parity of the HIP kernels with the FP64 oracle.

Each stage gets the same float32 input as the oracle stage; tolerances are
FP32 rounding levels.  Shapes: the small geometry and the first block of the
16-row clinical-FDK geometry.
"""
import numpy as np
import pytest
import torch

from listctrecons_amd import _lib, recon, scan
from oracle import oracle as O
from helpers import cpu_fan, oracle_plan, rel_rmse

pytestmark = pytest.mark.gpu

CASES = ["tiny", "helical_16x368_256x32"]


def block_window(w, pl, c, fan):
    s, _, _, _ = O.block_info(pl, c, 0)
    first = s + pl.data_offset
    return fan[first:first + pl.ctg.n_angle_add], s


@pytest.fixture(scope="module")
def data():
    out = {}
    for name in CASES:
        w = scan.workload(name)
        pl, a, c = oracle_plan(w)
        fan = cpu_fan(w)
        win, s = block_window(w, pl, c, fan)
        out[name] = (w, pl, a, c, fan, win, s)
    return out


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("kind,weight", [(0, 0), (1, 0), (2, 0), (2, 1), (1, 1)])
def test_rebin(ctx, data, name, kind, weight):
    w, pl, a, c, fan, win, s = data[name]
    g = pl.ctg
    print(f"rebin {name} kind={kind} weight={weight}", flush=True)
    ref = O.rebin(g, win, kind=kind)
    if weight:
        ref = O.weight_pi_ori(g, ref)
    d_fan = torch.from_numpy(np.ascontiguousarray(win)).cuda()
    d_par = torch.empty((g.n_angle, g.nt, g.ns), dtype=torch.float32, device="cuda")
    gg = _lib.CTGeom.from_buffer_copy(bytes(g))
    ctx.rebin(kind, weight, gg, d_fan, d_par)
    ctx.synchronize()
    got = d_par.cpu().numpy()
    assert rel_rmse(got, ref) < 1e-6
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())


@pytest.mark.parametrize("name", CASES)
def test_standalone_weighting(ctx, data, name):
    w, pl, a, c, fan, win, s = data[name]
    g = pl.ctg
    par = O.rebin(g, win, kind=2).astype(np.float32)
    ref = O.weight_pi_ori(g, par.astype(np.float64))
    d = torch.from_numpy(par).cuda()
    ctx.weighting(_lib.WEIGHT_PI_ORI, _lib.CTGeom.from_buffer_copy(bytes(g)), d)
    ctx.synchronize()
    assert rel_rmse(d.cpu().numpy(), ref) < 1e-6


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("window", [0, 1, 2])
def test_filter(ctx, data, name, window):
    w, pl, a, c, fan, win, s = data[name]
    g = pl.ctg
    print(f"filter {name} window={window}", flush=True)
    rows = O.rebin(g, win, kind=1).astype(np.float32)
    taps = recon.ramp_taps(window, g.ns, g.ds)
    ref = O.filter_rows(taps.astype(np.float64), rows.astype(np.float64))
    d = torch.from_numpy(rows).cuda()
    ctx.filter(_lib.CTGeom.from_buffer_copy(bytes(g)), taps, d, g.n_angle * g.nt)
    ctx.synchronize()
    assert rel_rmse(d.cpu().numpy(), ref) < 2e-6


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("algo,storage,tol", [(0, 0, 2e-6), (1, 0, 1e-4), (1, 1, 1e-3),
                                              (1, 2, 5e-3), (0, 1, 1e-3), (2, 0, 2e-6),
                                              (2, 1, 1e-3)])
def test_filter_into_backprojector_layout(ctx, data, name, algo, storage, tol):
    """Direct, BF16x3 and BF16x6 Toeplitz-MFMA filters, Q[view][ns+3][nt+8]
    output; the six-product form holds the direct FP32 filter's tolerance."""
    w, pl, a, c, fan, win, s = data[name]
    g = pl.ctg
    print(f"filter->Q {name} algo={algo} storage={storage}", flush=True)
    rows = O.rebin(g, win, kind=1).astype(np.float32)
    taps = recon.ramp_taps(2, g.ns, g.ds)
    ref = O.filter_rows(taps.astype(np.float64), rows.astype(np.float64))
    dt = {0: torch.float32, 1: torch.float16, 2: torch.bfloat16}[storage]
    q = torch.zeros((g.n_angle, g.ns + 3, g.nt + 8), dtype=dt, device="cuda")
    ctx.filter_to_bp_layout(algo, storage, g.ns, g.nt, taps, torch.from_numpy(rows).cuda(),
                            g.n_angle, q)
    ctx.synchronize()
    qn = q.float().cpu().numpy()
    got = qn[:, 1:g.ns + 1, 4:g.nt + 4].transpose(0, 2, 1)
    err = rel_rmse(got, ref)
    print(f"  rel rmse {err:.2e}", flush=True)
    assert err < tol
    # every guard slot stays zero
    mask = np.ones(qn.shape, bool)
    mask[:, 1:g.ns + 1, 4:g.nt + 4] = False
    assert not qn[mask].any()


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("kind", [0, 1])
def test_backproject(ctx, data, name, kind):
    w, pl, a, c, fan, win, s = data[name]
    g, ig = pl.ctg, pl.ig
    print(f"backproject {name} kind={kind}", flush=True)
    rk = 2 if kind == 1 else 1
    par = O.rebin(g, win, kind=rk)
    q = O.filter_rows(O.ramp_taps(g.ns, g.ds, w.window), par).astype(np.float32)
    betas = recon.block_betas(pl, a, s)
    ref = O.backproject(g, ig, q.astype(np.float64), betas, kind=kind)
    d_q = torch.from_numpy(q).cuda()
    d_img = torch.empty((32, ig.ny, ig.nx), dtype=torch.float32, device="cuda")
    ctx.backproject(kind, _lib.CTGeom.from_buffer_copy(bytes(g)),
                    _lib.ImgGeom.from_buffer_copy(bytes(ig)), d_q, betas, d_img)
    ctx.synchronize()
    got = d_img.cpu().numpy()
    err = rel_rmse(got, ref)
    print(f"  rel rmse {err:.3e}", flush=True)
    assert err < 2e-5


def test_projector_matches_oracle(ctx):
    w = scan.workload("helical_16x368_256x32")
    pl, a, c = oracle_plan(w)
    ell = w.ellipsoids()
    ref = O.project_ellipsoids(pl.ctg, a, c, ell, view0=100, n_views=40)
    out = torch.empty((40, w.nt, w.ns), dtype=torch.float32, device="cuda")
    ctx.project_ellipsoids(recon.scan_geometry(w), a, c, ell, out, view0=100, n_views=40)
    ctx.synchronize()
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-6, atol=1e-6)


def test_poisson_noise_statistics(ctx):
    n = 1 << 20
    for p_true, i0 in ((0.5, 1e5), (4.0, 2e5), (8.0, 1e5)):
        t = torch.full((n,), p_true, dtype=torch.float32, device="cuda")
        ctx.poisson_noise(t, i0, 7)
        ctx.synchronize()
        counts = i0 * np.exp(-t.cpu().numpy().astype(np.float64))
        lam = i0 * np.exp(-p_true)
        assert counts.mean() == pytest.approx(lam, rel=3e-3, abs=0.05)
        assert counts.var() == pytest.approx(lam, rel=0.02)
    # below 10 counts (multiplication method); zero counts are clipped to 1
    lam = 4.5
    t = torch.full((n,), float(np.log(1e5 / lam)), dtype=torch.float32, device="cuda")
    ctx.poisson_noise(t, 1e5, 11)
    ctx.synchronize()
    x = np.rint(1e5 * np.exp(-t.cpu().numpy().astype(np.float64)))
    p0 = np.exp(-lam)
    assert x.mean() == pytest.approx(lam + p0, rel=5e-3)
    assert x.var() == pytest.approx(lam + p0 - 2 * lam * p0 - p0 * p0, rel=0.02)
    # reproducible for a seed, different across seeds
    a = torch.full((1000,), 1.0, device="cuda")
    b = a.clone()
    c2 = a.clone()
    ctx.poisson_noise(a, 1e4, 1)
    ctx.poisson_noise(b, 1e4, 1)
    ctx.poisson_noise(c2, 1e4, 2)
    ctx.synchronize()
    assert torch.equal(a, b) and not torch.equal(a, c2)


@pytest.mark.parametrize("ns", [4, 257, 2048])
def test_filters_at_extreme_widths(ctx, ns):
    """The narrowest detector, a width that is not a multiple of the 256-lane
    row, and the widest the direct filter takes (2048 channels, 8 per lane):
    direct in place, direct into Q and (ns % 4 == 0) Toeplitz into Q."""
    rng = np.random.default_rng(ns)
    nt, views = 8, 2
    g = _lib.CTGeom()
    g.ns, g.nt, g.n_angle, g.ds = ns, nt, views, 0.5
    rows = rng.normal(size=(views * nt, ns)).astype(np.float32)
    taps = recon.ramp_taps(0, ns, 0.5)
    ref = O.filter_rows(taps.astype(np.float64), rows.astype(np.float64))
    print(f"filters ns={ns}", flush=True)
    d = torch.from_numpy(rows.copy()).cuda()
    ctx.filter(g, taps, d, views * nt)
    ctx.synchronize()
    assert rel_rmse(d.cpu().numpy(), ref) < 2e-6
    for algo, tol in ((0, 2e-6), (1, 1e-4)):
        if algo == 1 and ns % 4:
            continue
        q = torch.zeros((views, ns + 3, nt + 8), dtype=torch.float32, device="cuda")
        ctx.filter_to_bp_layout(algo, 0, ns, nt, taps, torch.from_numpy(rows).cuda(), views, q)
        ctx.synchronize()
        got = q.cpu().numpy()[:, 1:ns + 1, 4:nt + 4].transpose(0, 2, 1).reshape(views * nt, ns)
        assert rel_rmse(got, ref) < tol, algo
