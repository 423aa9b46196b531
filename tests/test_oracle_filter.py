"""Ramp taps and the row convolution (Src/Filter.cpp:8-20, CUDA/Filter.cu:10-50).

Pins: the Kak & Slaney closed form of the Ram-Lak taps, the windowed ramps'
frequency responses by numerical integration, the delta response, and an
independent numpy convolution.
"""
import math

import numpy as np
import pytest
from scipy import integrate

from listctrecons_amd import recon
from oracle import oracle as O


def test_ramlak_closed_form():
    ns, ds = 64, 0.57
    h = O.ramp_taps(ns, ds, 0)
    n = np.arange(2 * ns) - ns
    assert h[ns] == pytest.approx(1.0 / (4 * ds), rel=1e-15)
    odd = n % 2 != 0
    np.testing.assert_allclose(h[odd], -1.0 / (n[odd] ** 2 * math.pi ** 2 * ds), rtol=1e-12,
                               atol=1e-15)
    even = (n % 2 == 0) & (n != 0)
    assert np.max(np.abs(h[even])) < 1e-15  # sin(n pi) round-off only


def _response(n, ds, window):
    """tau * integral_{-1/2tau}^{1/2tau} |f| W(f) cos(2 pi f n tau) df."""
    fn = 1.0 / (2 * ds)
    if window == 1:
        W = lambda f: 0.5 * (1 + math.cos(math.pi * f / fn))
    elif window == 2:
        W = lambda f: 1.0 if f == 0 else math.sin(math.pi * f * ds) / (math.pi * f * ds)
    else:
        W = lambda f: 1.0
    val, _ = integrate.quad(lambda f: 2 * f * W(f) * math.cos(2 * math.pi * f * n * ds), 0, fn,
                            limit=200)
    return ds * val


@pytest.mark.parametrize("window", [0, 1, 2])
def test_windowed_taps_match_frequency_response(window):
    ns, ds = 32, 0.8
    h = O.ramp_taps(ns, ds, window)
    for n in (-7, -2, -1, 0, 1, 2, 3, 10):
        assert h[ns + n] == pytest.approx(_response(n, ds, window), rel=1e-9, abs=1e-12)


def test_windows_attenuate_high_frequency():
    ns, ds = 256, 0.57
    f_nyq = np.cos(np.pi * (np.arange(2 * ns) - ns))  # (-1)^n
    r = [np.sum(O.ramp_taps(ns, ds, w) * f_nyq) for w in (0, 1, 2)]
    # Ram-Lak passes Nyquist at ~1/(2 ds); Hann kills it; Shepp-Logan keeps 2/pi
    assert r[0] == pytest.approx(1 / (2 * ds), rel=2e-3)
    assert abs(r[1]) < 2e-3 * r[0]
    assert r[2] == pytest.approx(2 / math.pi / (2 * ds), rel=5e-3)


def test_product_taps_are_oracle_taps_in_float32():
    for window in (0, 1, 2):
        for ns, ds in ((368, 0.570074), (736, 0.570074), (64, 2.0)):
            hp = recon.ramp_taps(window, ns, ds)
            ho = O.ramp_taps(ns, float(np.float32(ds)), window)
            np.testing.assert_array_equal(hp, ho.astype(np.float32))


def test_filter_delta_response_is_the_taps():
    ns = 48
    h = O.ramp_taps(ns, 0.57, 0)
    rows = np.zeros((3, ns))
    for r, k0 in enumerate((0, 17, ns - 1)):
        rows[r, k0] = 1.0
    out = O.filter_rows(h, rows)
    for r, k0 in enumerate((0, 17, ns - 1)):
        np.testing.assert_array_equal(out[r], h[np.arange(ns) - k0 + ns])


def test_filter_matches_numpy_convolution():
    rng = np.random.default_rng(0)
    ns = 100
    h = O.ramp_taps(ns, 0.57, 2)
    rows = rng.normal(size=(5, ns))
    out = O.filter_rows(h, rows)
    for r in range(5):
        full = np.convolve(rows[r], h)  # full[m] = sum_k row[k] h[m-k]
        np.testing.assert_allclose(out[r], full[ns:2 * ns], rtol=1e-12, atol=1e-12)
