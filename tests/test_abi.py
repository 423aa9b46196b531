"""The C ABI library: it loads, exports every function include/lct.h declares,
and its host-only functions (planner, block windows, taps) run without a GPU.
No kernel is launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from listctrecons_amd import _lib, recon, scan

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lct_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = declared("lct.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes table covers all of them
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_version_and_errors():
    lib = _lib.lib()
    assert lib.lct_version().decode().startswith("lct ")
    rc = lib.lct_ramp_taps(0, 0, C.c_float(1.0), None)
    assert rc == _lib.LCT_ERR_INVALID
    assert b"lct_ramp_taps" in lib.lct_last_error()
    with pytest.raises(_lib.LctError):
        recon.ramp_taps(7, 16, 1.0)


def test_host_planner_without_gpu():
    w = scan.workload("helical_64x736_512x256")
    a, c = w.angles_couch()
    pl = recon.plan_helical(w.prm(), a, c)
    assert (pl.ctg.add_projections, pl.ctg.n_angle_add, pl.ctg.n_angle) == (81, 2148, 1986)
    assert pl.n_blocks == 8
    s0 = recon.block_info(pl, c, 0)[0]
    s1 = recon.block_info(pl, c, 1)[0]
    # consecutive blocks advance by 32 slices of table travel
    assert (s1 - s0) == pytest.approx(32 * w.dz / (w.pitch_value / w.n_proj_turn), abs=1.0)


@pytest.mark.parametrize("data_offset", [0, 7])
def test_slab_views_cover_the_block_windows(data_offset):
    """lct_slab_views == union of the oracle's BlockInfo windows + DataOffset."""
    from oracle import oracle as O
    from helpers import oracle_plan
    w = scan.workload("helical_64x736_512x256", data_offset=data_offset)
    opl, a, c = oracle_plan(w)
    pl = recon.plan_helical(w.prm(), a, c)
    nb = pl.n_blocks
    for b0, b1 in ((0, 1), (3, 4), (2, 6), (0, nb)):
        starts = [O.block_info(opl, c, b)[0] for b in range(b0, b1)]
        want = (min(starts) + data_offset, max(starts) + data_offset + pl.ctg.n_angle_add)
        assert recon.slab_views(pl, c, b0, b1) == want
    for b0, b1 in ((-1, 2), (3, 3), (0, nb + 1)):
        with pytest.raises(_lib.LctError) as e:
            recon.slab_views(pl, c, b0, b1)
        assert e.value.code == _lib.LCT_ERR_INVALID


def test_planner_range_error_code():
    w = scan.workload("tiny", readings=300)
    a, c = w.angles_couch()
    with pytest.raises(_lib.LctError) as e:
        recon.plan_helical(w.prm(), a, c)
    assert e.value.code == _lib.LCT_ERR_RANGE
    assert "outside of allowed data range" in str(e.value)


def test_ctx_create_fails_cleanly_without_gpu():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except Exception:
        pass
    h = C.c_void_p()
    rc = _lib.lib().lct_ctx_create(0, C.byref(h))
    assert rc < 0 and not h.value
