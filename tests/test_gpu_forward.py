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
