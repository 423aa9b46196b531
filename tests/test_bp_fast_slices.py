"""The backprojector's fast-slice rule (bp_windows_kernel in
listctrecons_amd/csrc/lct_resident.hip): a slice is marked fast for a
(tile, view pair) when both rays' detector row offsets qh(z) = z K + O lie
within wt - 1/4 - e at the tile's four corner columns, e = 2 (dx + dy) from
the departures dx, dy of qh at two edge midpoints from the corners' bilinear
interpolation.  On a fast slice the kernel skips the detector tests, so the
rule must never mark a slice on which any column of the tile has a ray off
the detector (|qh| > wt, Backprojector.cu:66-83).

Checked here by brute force over every column of sampled tiles, on the
benchmark workloads and the seeded random geometries of the GPU parity
tests, with the rule restated in float64.  A rule that looks at too small a
tile (a planted mistake) is caught.
"""
import math

import numpy as np
import pytest

from listctrecons_amd import recon, scan
from test_gpu_random_geometry import random_workload

TILE_X, TILE_Y = 16, 32


def _geometry(w):
    a, c = w.angles_couch()
    pl = recon.plan_helical(w.prm(), a, c)
    g, ig = pl.ctg, pl.ig
    h = g.n_proj_turn // 2
    dzv = g.pitch / g.n_proj_turn
    amid = []
    for iz in range(32):
        zc = (iz - ig.wz) * ig.dz
        am = int(round((zc - g.source_z0) / dzv))
        amid.append(min(max(am, 0), g.n_angle - 1))
    s0 = recon.block_info(pl, c, 0)[0]
    betas = recon.block_betas(pl, a, s0)
    return g, ig, h, dzv, amid, betas


def _lines(g, ig, h, dzv, betas, x, y, j):
    """K, O of qh(z) = z K + O for both rays of pair j through columns (x, y)."""
    kc = g.dso * g.z_dir / g.dt
    xc = (np.asarray(x, float) - ig.wx) * ig.dx
    yc = (np.asarray(y, float) - ig.wy) * ig.dy
    b1, b2 = betas[j], betas[j + h]
    phat = -xc * math.cos(b1) - yc * math.sin(b1)
    theta = np.arcsin(phat / g.dso)
    root = np.sqrt(g.dso ** 2 - phat ** 2)
    lhat = root + xc * math.sin(b1) - yc * math.cos(b1)
    lhat2 = root + xc * math.sin(b2) - yc * math.cos(b2)
    zr = g.source_z0 + (j + theta / g.orbit) * dzv
    zr2 = g.source_z0 + (j + h - theta / g.orbit) * dzv
    k1, k2 = kc / lhat, kc / lhat2
    K = np.array([k1 * ig.dz, k2 * ig.dz])
    return K, np.array([-ig.wz * K[0] - zr * k1, -ig.wz * K[1] - zr2 * k2])


def fast_mask(geo, x0, x1, y0, y1, j):
    g, ig, h, dzv, _, betas = geo
    corners = [_lines(g, ig, h, dzv, betas, x, y, j) for x in (x0, x1) for y in (y0, y1)]
    (K00, O00), (K01, O01), (K10, O10), _ = corners
    Kx, Ox = _lines(g, ig, h, dzv, betas, 0.5 * (x0 + x1), y0, j)
    Ky, Oy = _lines(g, ig, h, dzv, betas, x0, 0.5 * (y0 + y1), j)
    e = 0.0
    for v in range(2):
        dx = max(abs(z * Kx[v] + Ox[v] - 0.5 * (z * K00[v] + O00[v] + z * K10[v] + O10[v]))
                 for z in (0.0, 31.0))
        dy = max(abs(z * Ky[v] + Oy[v] - 0.5 * (z * K00[v] + O00[v] + z * K01[v] + O01[v]))
                 for z in (0.0, 31.0))
        e = max(e, 2.0 * (dx + dy))
    L = g.wt - 0.25 - e
    if L <= 0:
        return 0
    zlo, zhi = -1e30, 1e30
    for K, O in corners:
        for v in range(2):
            za, zb = (-L - O[v]) / K[v], (L - O[v]) / K[v]
            zlo, zhi = max(zlo, min(za, zb)), min(zhi, max(za, zb))
    z0 = math.ceil(min(max(zlo + 1e-3, 0.0), 32.0))
    z1 = math.floor(max(min(zhi - 1e-3, 31.0), -1.0))
    return sum(1 << z for z in range(z0, z1 + 1))


def violations(w, trials, seed, tile=(TILE_X, TILE_Y)):
    """(fast slices checked, fast slices with a column's ray off the detector)."""
    geo = _geometry(w)
    g, ig, h, dzv, amid, betas = geo
    rng = np.random.default_rng(seed)
    ntx, nty = (ig.nx + TILE_X - 1) // TILE_X, (ig.ny + TILE_Y - 1) // TILE_Y
    checked = bad = 0
    for _ in range(trials):
        j = int(rng.integers(amid[0] - h, amid[31]))
        if j < 0 or j + h >= len(betas):
            continue
        x0, y0 = int(rng.integers(0, ntx)) * TILE_X, int(rng.integers(0, nty)) * TILE_Y
        fm = fast_mask(geo, x0, x0 + tile[0] - 1, y0, y0 + tile[1] - 1, j)
        ix = x0 + np.arange(TILE_X * TILE_Y) % TILE_X
        iy = y0 + np.arange(TILE_X * TILE_Y) // TILE_X
        K, O = _lines(g, ig, h, dzv, betas, ix, iy, j)
        for z in range(32):
            if not (amid[z] - h <= j < amid[z]) or not (fm >> z) & 1:
                continue
            checked += 1
            if max(np.abs(z * K[0] + O[0]).max(), np.abs(z * K[1] + O[1]).max()) > g.wt:
                bad += 1
    return checked, bad


@pytest.mark.parametrize("name", ["helical_16x368_256x32", "helical_64x736_512x256",
                                  "helical_96x896_1024x512"])
def test_fast_slices_keep_every_ray_on_the_detector(name):
    checked, bad = violations(scan.workload(name), trials=150, seed=1)
    assert checked > 500 and bad == 0


@pytest.mark.parametrize("seed", range(12))
def test_fast_slices_on_random_geometries(seed):
    checked, bad = violations(random_workload(seed), trials=60, seed=seed)
    assert bad == 0


def test_a_rule_on_too_small_a_tile_is_caught():
    """Corners of the first 4 x 8 columns only: slices where the rest of the
    tile leaves the detector get marked."""
    bad = sum(violations(random_workload(s), trials=60, seed=s, tile=(4, 8))[1] for s in range(12))
    assert bad > 0
