"""Forward projector (reprojection of voxel volumes) pins: a constant volume
reprojects to the chord length inside the voxel-centre box (trilinear
interpolation of a constant is that constant), and a voxelised ball
reprojects close to its analytic line integrals."""
import math

import numpy as np
import pytest
import torch

from listctrecons_amd import _lib, recon, scan
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def setup(nx=40, nz=24, dx=2.0, dz=2.0):
    w = scan.workload("tiny")
    g = recon.scan_geometry(w)
    ig = _lib.ImgGeom()
    ig.nx = ig.ny = nx
    ig.nz = nz
    ig.dx = ig.dy = dx
    ig.dz = dz
    ig.wx = ig.wy = (nx - 1) / 2.0
    a, c = w.angles_couch()
    return w, g, ig, a, c


def box_chords(g, a, c, views, lo, hi):
    out = np.zeros((len(views), g.nt, g.ns))
    for vi, v in enumerate(views):
        b = float(a[v])
        S = np.array([-g.dso * math.sin(b), g.dso * math.cos(b), float(c[v])])
        for it in range(g.nt):
            for is_ in range(g.ns):
                al = (is_ - g.ws) * g.ds / g.dso
                sl = (it - g.wt) * g.dt / g.dso
                d = np.array([math.sin(b - al), -math.cos(b - al), sl])
                d /= np.linalg.norm(d)
                t0, t1 = -1e30, 1e30
                for k in range(3):
                    ta, tb = (lo[k] - S[k]) / d[k], (hi[k] - S[k]) / d[k]
                    t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
                out[vi, it, is_] = max(t1 - t0, 0.0)
    return out


def test_constant_volume_gives_box_chords(ctx):
    w, g, ig, a, c = setup()
    zmid = float(c[500])
    z0 = zmid - (ig.nz - 1) * ig.dz / 2
    vol = torch.full((ig.nz, ig.ny, ig.nx), 0.5, device="cuda")
    views = list(range(480, 520))
    out = torch.empty((len(views), g.nt, g.ns), device="cuda")
    ctx.forward_project(g, a, c, vol, ig, z0, out, view0=views[0], n_views=len(views), step=0.05)
    ctx.synchronize()
    lo = [-ig.wx * ig.dx, -ig.wy * ig.dy, z0]
    hi = [(ig.nx - 1 - ig.wx) * ig.dx, (ig.ny - 1 - ig.wy) * ig.dy, z0 + (ig.nz - 1) * ig.dz]
    want = 0.5 * box_chords(g, a, c, views, lo, hi)
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=2e-4, atol=2e-4)


def test_voxelised_ball_reprojects_like_the_ball(ctx):
    w, g, ig, a, c = setup(nx=64, nz=48, dx=1.0, dz=1.0)
    zmid = float(c[500])
    z0 = zmid - (ig.nz - 1) * ig.dz / 2
    ell = scan.make_ellipsoid([3.0, -2.0, zmid], [15.0, 15.0, 15.0], np.eye(3), 0.02)[None]
    oig = O.ImgGeom.from_buffer_copy(bytes(ig))
    vol = O.voxelize(oig, z0, ig.nz, ell, 4).astype(np.float32)
    views = list(range(490, 510))
    out = torch.empty((len(views), g.nt, g.ns), device="cuda")
    ctx.forward_project(g, a, c, torch.from_numpy(vol).cuda(), ig, z0, out, view0=views[0],
                        n_views=len(views), step=0.25)
    ctx.synchronize()
    og = O.CTGeom.from_buffer_copy(bytes(g))
    ref = O.project_ellipsoids(og, a, c, ell, view0=views[0], n_views=len(views))
    got = out.cpu().numpy()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 0.02  # voxelisation of the ball surface at 1 mm


def gpu_forward(ctx, g, ig, z0, vol, ang, zs, step):
    cg = _lib.CTGeom.from_buffer_copy(bytes(g))  # same layout as the oracle's
    vg = _lib.ImgGeom.from_buffer_copy(bytes(ig))
    out = torch.empty((ang.size, g.nt, g.ns), dtype=torch.float32, device="cuda")
    ctx.forward_project(cg, ang, zs, torch.from_numpy(np.ascontiguousarray(vol, np.float32)).cuda(),
                        vg, z0, out, view0=0, n_views=ang.size, step=step)
    ctx.synchronize()
    return out.cpu().numpy().astype(np.float64)


def test_linear_volume_matches_the_closed_form_and_the_oracle(ctx):
    """Same closed form as the oracle's pin (ray sum = chord x midpoint value
    for a linear volume), here through the FP32 kernel."""
    from test_oracle_phantom import linear_case
    g, ig, z0, vol, f, lo, hi, ang, zs = linear_case()
    got = gpu_forward(ctx, g, ig, z0, vol, ang, zs, 1.7)
    ref = O.forward_project(g, ang, zs, vol, ig, z0, step=1.7)
    assert np.abs(ref).max() > 1.0
    np.testing.assert_allclose(got, ref, rtol=2e-5, atol=2e-5 * np.abs(ref).max())


def test_random_volume_matches_the_oracle(ctx):
    """A smooth random volume (trilinear interpolation and the midpoint rule
    no longer exact) against the FP64 restatement with the same steps."""
    from test_oracle_phantom import linear_case
    g, ig, z0, _, _, _, _, ang, zs = linear_case()
    rng = np.random.default_rng(3)
    vol = rng.uniform(0.0, 1.0, (ig.nz, ig.ny, ig.nx))
    got = gpu_forward(ctx, g, ig, z0, vol, ang, zs, 0.9)
    ref = O.forward_project(g, ang, zs, vol, ig, z0, step=0.9)
    # FP32 sums; a ray whose chord / step sits within rounding of an integer
    # may take one step more or fewer than the FP64 restatement
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-4


def gpu_adjoint(ctx, g, ig, z0, proj, ang, zs, step, nz):
    cg = _lib.CTGeom.from_buffer_copy(bytes(g))
    vg = _lib.ImgGeom.from_buffer_copy(bytes(ig))
    out = torch.full((nz, ig.ny, ig.nx), 7.0, dtype=torch.float32, device="cuda")  # overwritten
    p = torch.from_numpy(np.ascontiguousarray(proj, np.float32)).cuda()
    ctx.forward_project_adjoint(cg, ang, zs, p, vg, z0, out, view0=0, step=step)
    ctx.synchronize()
    return out.cpu().numpy().astype(np.float64)


def test_adjoint_matches_the_oracle_and_pairs_with_the_projector(ctx):
    """The FP32 adjoint against the FP64 transpose (itself pinned by the
    explicit matrix), the dot-product identity of the GPU pair, and a second,
    smaller call that must overwrite rather than accumulate."""
    from test_oracle_phantom import adjoint_case
    g, ig, z0, ang, zs = adjoint_case(nx=20, ny=18, nz=12)
    rng = np.random.default_rng(7)
    y = rng.standard_normal((ang.size, g.nt, g.ns)).astype(np.float32)
    x = rng.standard_normal((ig.nz, ig.ny, ig.nx)).astype(np.float32)
    got = gpu_adjoint(ctx, g, ig, z0, y, ang, zs, 1.3, ig.nz)
    ref = O.forward_project_adjoint(g, ang, zs, y, ig, z0, ig.nz, step=1.3)
    assert np.abs(ref).max() > 1.0
    # as for the projector: a ray whose chord / step sits within rounding of
    # an integer may take one step more or fewer than the FP64 restatement
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-4
    ax = gpu_forward(ctx, g, ig, z0, x, ang, zs, 1.3)
    lhs = float((ax * y).sum())
    rhs = float((x.astype(np.float64) * got).sum())
    assert abs(lhs) > 1.0
    assert lhs == pytest.approx(rhs, rel=1e-5)
    got2 = gpu_adjoint(ctx, g, ig, z0, y[:2], ang[:2], zs[:2], 1.3, ig.nz)
    ref2 = O.forward_project_adjoint(g, ang[:2], zs[:2], y[:2], ig, z0, ig.nz, step=1.3)
    assert np.linalg.norm(got2 - ref2) / np.linalg.norm(ref2) < 1e-4


def test_projector_wrappers_refuse_mismatched_buffers(ctx):
    from test_oracle_phantom import adjoint_case
    g, ig, z0, ang, zs = adjoint_case()
    cg = _lib.CTGeom.from_buffer_copy(bytes(g))
    vg = _lib.ImgGeom.from_buffer_copy(bytes(ig))
    vol = torch.zeros((ig.nz, ig.ny + 1, ig.nx), device="cuda")
    out = torch.zeros((ang.size, g.nt, g.ns), device="cuda")
    with pytest.raises(ValueError):
        ctx.forward_project(cg, ang, zs, vol, vg, z0, out, step=1.0)
    with pytest.raises(ValueError):
        ctx.forward_project_adjoint(cg, ang, zs, out, vg, z0, vol, step=1.0)
    with pytest.raises(ValueError):
        ctx.forward_project(cg, ang, zs, vol[:, :ig.ny], vg, z0, out.double(), step=1.0)
