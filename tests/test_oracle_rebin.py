"""Rebin and pre-weighting pins (CUDA/Rebiner.cu, CUDA/PreWeighting.cu).

Pins: the null rebin is an exact view shift; with ds / dso equal to the orbit
step the fan-to-parallel offsets are whole views, so the rebin must be an
exact copy; a linear-in-view input is reproduced exactly by the view-axis
interpolation; constants stay constant away from the border; the PI-original
trilinear sampling reproduces linear functions of the three texel
coordinates; the cone weight is 1 on the central ray and never above 1.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def geom(ns=16, nt=3, ws=7.0, dso=100.0, orbit=0.01, ds=1.0, add=10, na=40, pitch=5.0,
         npt=628):
    g = O.CTGeom()
    g.orbit, g.orbit_start = orbit, 0.0
    g.ns, g.nt = ns, nt
    g.ws, g.wt = ws, (nt - 1) / 2.0
    g.n_angle, g.n_angle_add = na, na + 2 * add
    g.ds, g.dt, g.pitch, g.source_z0, g.dso = ds, ds, pitch, -pitch / 2, dso
    g.n_proj_turn, g.z_dir, g.add_projections = npt, 1, add
    return g


def test_null_rebin_is_a_view_shift():
    g = geom()
    fan = np.random.default_rng(0).normal(size=(g.n_angle_add, g.nt, g.ns)).astype(np.float32)
    par = O.rebin(g, fan, kind=0)
    np.testing.assert_array_equal(par, fan[g.add_projections:g.add_projections + g.n_angle])


def test_fan2par_whole_view_offsets_copy_exactly():
    # alpha / orbit = (is - ws) ds / dso / orbit = is - ws  (whole views)
    g = geom(ds=float(np.float32(100.0 * 0.01)), orbit=float(np.float32(0.01)))
    assert abs(g.ds / g.dso / g.orbit - 1.0) < 1e-6
    fan = np.random.default_rng(1).normal(size=(g.n_angle_add, g.nt, g.ns)).astype(np.float32)
    par = O.rebin(g, fan, kind=1)
    for ia in range(g.n_angle):
        for is_ in range(g.ns):
            v = ia + g.add_projections + int(round(is_ - g.ws))
            np.testing.assert_allclose(par[ia, :, is_], fan[v, :, is_], rtol=0, atol=2e-6)


def test_fan2par_linear_in_view_is_exact():
    g = geom(orbit=0.013)
    v = np.arange(g.n_angle_add, dtype=np.float64)
    fan = (0.25 * v[:, None, None] + np.arange(g.nt)[None, :, None]
           + 0 * np.arange(g.ns)[None, None, :]).astype(np.float32)
    par = O.rebin(g, fan, kind=1)
    is_ = np.arange(g.ns)
    vcoord = (np.arange(g.n_angle)[:, None] + g.add_projections
              + (is_[None, :] - g.ws) * g.ds / g.dso / g.orbit)
    expect = 0.25 * vcoord[:, None, :] + np.arange(g.nt)[None, :, None]
    inside = (vcoord >= 0) & (vcoord <= g.n_angle_add - 1)
    assert inside.all()
    np.testing.assert_allclose(par, expect, rtol=1e-12, atol=1e-9)


def test_fan2par_border_reads_zero():
    g = geom(orbit=0.001, add=2)  # offsets of up to 9 ds/dso/orbit = 90 views: past the window
    fan = np.ones((g.n_angle_add, g.nt, g.ns), np.float32)
    par = O.rebin(g, fan, kind=1)
    assert par.min() == 0.0 and par.max() == 1.0


def test_pi_ori_trilinear_reproduces_linear_functions():
    g = geom(ns=32, nt=8, ws=15.5, dso=60.0, ds=0.5, orbit=0.02, add=12, na=30, pitch=3.0)
    fan = np.zeros((g.n_angle_add, g.nt, g.ns), np.float32)
    z, y, x = np.meshgrid(np.arange(g.n_angle_add), np.arange(g.nt), np.arange(g.ns),
                          indexing="ij")
    fan[:] = 0.01 * z + 0.5 * y + 0.125 * x
    par = O.rebin(g, fan, kind=2)
    for ia in (0, 7, 29):
        for it in range(g.nt):
            for is_ in range(g.ns):
                s = (is_ - g.ws) * g.ds
                gam = math.asin(s / g.dso)
                pI = gam * g.dso / g.ds + g.ws
                q = ((it - g.wt) * g.pitch / (2 * (g.nt - 1)) - gam * g.pitch / (2 * math.pi)) \
                    / math.cos(gam)
                qI = q / g.dt + g.wt
                bI = ia + gam / g.orbit + g.add_projections
                if 0 <= pI <= g.ns - 1 and 0 <= qI <= g.nt - 1 and 0 <= bI <= g.n_angle_add - 1:
                    assert par[ia, it, is_] == pytest.approx(0.01 * bI + 0.5 * qI + 0.125 * pI,
                                                             rel=1e-6, abs=1e-6)


def test_pi_ori_weight():
    g = geom(ns=33, nt=9, ws=16.0, dso=60.0, ds=0.5, pitch=3.0)
    w = O.weight_pi_ori(g, np.ones((g.n_angle, g.nt, g.ns)))
    assert w[0, 4, 16] == 1.0  # s = 0 and t = 0
    assert (w <= 1.0).all() and (w > 0.9).all()
    s = (np.arange(g.ns) - g.ws) * g.ds
    t = (np.arange(g.nt) - g.wt) * g.pitch / (2 * g.nt)
    gam = np.arcsin(s / g.dso)
    a = g.dso ** 2 - s ** 2
    expect = np.sqrt(a)[None, :] / np.sqrt(a[None, :] + (t[:, None] - gam[None, :] * g.pitch
                                                        / (2 * math.pi)) ** 2)
    np.testing.assert_allclose(w[3], expect, rtol=1e-14)


def test_pi_ori_weight_is_the_cone_cosine_of_the_rebinned_ray():
    """Pins lco_weight_pi_ori by the geometry it encodes rather than by its
    formula.  A parallel-rebinned ray at channel offset s comes from a source
    at horizontal distance D = sqrt(dso^2 - s^2) from the virtual detector
    through the axis, displaced along z by gamma P / (2 pi) with
    gamma = asin(s / dso) (the helix advances P per turn, the fan angle gamma
    is gamma / (2 pi) of a turn).  The weight is the cosine of the ray's cone
    angle towards detector row height t: w = D / sqrt(D^2 + (t - z_src)^2).

    From the weights alone, per channel: 1/w^2 - 1 = (t - z_src)^2 / D^2 is a
    quadratic in the row height t, so a fit through the rows recovers the
    source's z (vertex) and D (curvature), which must match the geometry; at
    the central channel the recovered row heights are spaced P / (2 nt)
    (CUDA/PreWeighting.cu:8-29 scales the rows by 2 nt; the PI-original
    rebin, Rebiner.cu:156-179, by 2 (nt - 1)).  A flipped sign of the gamma
    term, a linear gamma = s / dso, a dropped s^2 in D, a 2 pi for 2 pi / P
    or rows scaled by 2 (nt - 1) each fail one of these checks."""
    g = geom(ns=41, nt=12, ws=20.0, dso=60.0, ds=1.2, pitch=24.0)
    w = O.weight_pi_ori(g, np.ones((g.n_angle, g.nt, g.ns)))[0]  # [nt][ns]
    row_step = g.pitch / (2 * g.nt)
    # central channel (s = 0, gamma = 0): 1/w^2 - 1 = t^2 / dso^2
    t_abs = g.dso * np.sqrt(1.0 / w[:, 20] ** 2 - 1.0)
    t_signed = np.where(np.arange(g.nt) < g.wt, -t_abs, t_abs)
    np.testing.assert_allclose(np.diff(t_signed), row_step, rtol=1e-9)
    assert abs(row_step - g.pitch / (2 * (g.nt - 1))) > 1e-3  # the two scalings differ here
    t = (np.arange(g.nt) - g.wt) * row_step
    for is_ in (0, 5, 13, 27, 40):
        s = (is_ - g.ws) * g.ds
        a, b, c = np.polyfit(t, 1.0 / w[:, is_] ** 2 - 1.0, 2)
        z_src, D = -b / (2 * a), 1.0 / math.sqrt(a)
        assert D == pytest.approx(math.sqrt(g.dso ** 2 - s ** 2), rel=1e-9)
        assert z_src == pytest.approx(math.asin(s / g.dso) * g.pitch / (2 * math.pi),
                                      rel=1e-8, abs=1e-10)
        assert c == pytest.approx(z_src ** 2 / D ** 2, rel=1e-7, abs=1e-12)
