"""Planner pins (Src/DataStructs.cpp:12-34, 77-108, 121-250).

The expected counts below were derived by hand from the reference formulas
(guard views add = int(max(cc, ns-cc) ds npt / (2 dso pi) + 10), padding
P (npt/2 + add + 256) / npt, 32-slice rounding, NA+ = span + 100,
NA = NA+ - 2 add) independently of the oracle code, and the product planner
in liblct.so must agree with the oracle field for field.
"""
import math

import numpy as np
import pytest

from listctrecons_amd import recon, scan
from oracle import oracle as O


def oracle_prm(w):
    p = O.Prm()
    pp = w.prm()
    for f, _ in O.Prm._fields_:
        setattr(p, f, getattr(pp, f))
    return p


EXPECTED = {
    # name: (add, NA+, NA, slices, blocks)
    "helical_16x368_256x32": (27, 2296, 2242, 32, 1),
    "helical_64x736_512x256": (81, 2148, 1986, 256, 8),
    "helical_96x896_1024x512": (183, 3292, 2926, 512, 16),
}


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_planner_counts(name):
    w = scan.workload(name)
    a, c = w.angles_couch()
    pl = O.plan_helical(oracle_prm(w), a, c)
    add, naa, na, slices, blocks = EXPECTED[name]
    assert pl.ctg.add_projections == add
    assert pl.ctg.n_angle_add == naa
    assert pl.ctg.n_angle == na
    assert pl.n_slices_recon == slices
    assert pl.n_blocks == blocks
    assert pl.ctg.z_dir == 1
    assert pl.ig.nz == 32 and pl.ig.wz == 0.0 and pl.ig.dy == pl.ig.dx


def test_guard_views_closed_form():
    for name in EXPECTED:
        w = scan.workload(name)
        half = max(w.central_channel, w.ns - w.central_channel)
        add = int(half * w.ds * w.n_proj_turn / (2 * w.dso * math.pi) + 10)
        a, c = w.angles_couch()
        assert O.plan_helical(oracle_prm(w), a, c).ctg.add_projections == add


def test_two_turn_scans_are_rejected():
    # the prose's 2-rotation scan is shorter than the planner's padding
    w = scan.workload("helical_16x368_256x32", readings=1152)
    a, c = w.angles_couch()
    with pytest.raises(ValueError):
        O.plan_helical(oracle_prm(w), a, c)
    with pytest.raises(Exception):
        recon.plan_helical(w.prm(), a, c)


def test_readings_range_for_256_slices():
    # Readings in [7979, 8728] give exactly 256 slices for the clinical scan
    for r, slices in ((7978, 224), (7979, 256), (8728, 256), (8729, 288)):
        w = scan.workload("helical_64x736_512x256", readings=r)
        a, c = w.angles_couch()
        assert O.plan_helical(oracle_prm(w), a, c).n_slices_recon == slices


VARIANTS = [
    ("tiny", {}),
    ("tiny", {"z_dir": -1}),
    ("tiny", {"view_offset": 3, "data_offset": 5}),
    ("tiny", {"offset_x": 1.5, "offset_y": -2.0}),
    ("tiny", {"angle0": 6.2}),  # wrap between views 0 and 1 -> orbit from views 1, 2
    ("helical_16x368_256x32", {}),
    ("helical_64x736_512x256", {}),
    ("helical_16x368_p15_256x128", {}),
]


@pytest.mark.parametrize("name,kw", VARIANTS)
def test_product_planner_matches_oracle(name, kw):
    w = scan.workload(name, **kw)
    a, c = w.angles_couch()
    p1 = O.plan_helical(oracle_prm(w), a, c)
    p2 = recon.plan_helical(w.prm(), a, c)
    assert bytes(p1) == bytes(p2)
    for b in range(p1.n_blocks):
        s, e, bs, be = O.block_info(p1, c, b)
        s2, e2, bs2, be2, ref = recon.block_info(p2, c, b)
        assert (s, e, bs, be) == (s2, e2, bs2, be2)
        assert ref == s + p1.ctg.n_proj_turn // 2 + p1.ctg.add_projections


def test_descending_table():
    w = scan.workload("tiny", z_dir=-1)
    a, c = w.angles_couch()
    pl = O.plan_helical(oracle_prm(w), a, c)
    assert pl.ctg.z_dir == -1 and pl.recon_direction == -1
    assert pl.recon_end_pos < pl.recon_start_pos


def test_orbit_wrap_rule():
    w = scan.workload("tiny", angle0=2 * math.pi - 0.01)
    a, c = w.angles_couch()
    assert a[0] > 6 and a[1] < 1
    pl = O.plan_helical(oracle_prm(w), a, c)
    assert pl.ctg.orbit == np.float32(a[2] - a[1])


def test_array_search():
    asc = np.array([0.0, 1.0, 2.0, 3.0], np.float32)
    assert O.array_search(1.5, asc, 1) == 2
    assert O.array_search(-1.0, asc, 1) == 0
    assert O.array_search(9.0, asc, 1) == 4  # stops at n (the reference reads a[n])
    desc = asc[::-1].copy()
    assert O.array_search(1.5, desc, -1) == 2
    assert O.array_search(2.0, asc, 0) == 2
