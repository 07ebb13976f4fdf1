"""Known answers that pin the frame oracle (oracle/frame_oracle.py).

Each value below is derived by hand from the reference formulas
(turbo_wrapper.py:317-337, turbo_main.py:830-890, 2109-2113, 2154-2214),
chosen so that a plausible slip -- rounding instead of truncating, round
half up instead of half to even, a 255 instead of 254 scale, a wrong
display size, a filter that skips its first-call seeding, an error term
updated inside the dead band -- changes the result.
"""
import math

import numpy as np
import pytest

from oracle import frame_oracle as FO


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_pixels_known_answers(dtype):
    # 1 + norm * 254, truncated: 0 -> 1, 0.25 -> 64 (64.5), 0.5 -> 128, 0.75 -> 191 (191.5), 1 -> 255
    f = np.array([[0.0, 0.25, 0.5], [0.75, 1.0, 0.0]], dtype=dtype)
    assert FO.float_to_pixels(f).tolist() == [[1, 64, 128], [191, 255, 1]]
    # an affine change of the plane leaves the pixels alone
    assert FO.float_to_pixels(f * dtype(4) - dtype(2)).tolist() == [[1, 64, 128], [191, 255, 1]]


def test_pixels_constant_plane_is_mid_grey():
    assert (FO.float_to_pixels(np.full((3, 5), 7.5, np.float32)) == 128).all()
    # a range under float32(1e-12) counts as constant
    f = np.array([0.0, 5e-13], dtype=np.float64)
    assert FO.float_to_pixels(f).tolist() == [128, 128]


def test_pixel_stats_two_level():
    p = np.array([100, 110] * 8, dtype=np.uint8)
    mu, sig = FO.pixel_stats(p)
    assert mu == 105.0 and sig == 5.0
    assert FO.pixel_stats(np.full(9, 128, np.uint8)) == (128.0, 0.0)


def test_stretch_known_answers():
    # mu 105, sigma 5: lo 92.5, inv float32(10.2); 7.5 * 10.2 and 17.5 * 10.2 land on
    # 76.5 and 178.5 in float32 and round half to even -> 76, 178 (half up: 77, 179)
    p = np.array([[100, 110], [110, 100]], dtype=np.uint8)
    assert FO.display_stretch(p, 105.0, 5.0).tolist() == [[76, 178], [178, 76]]
    # the band edges map to 0 and 255 and clip beyond
    q = np.array([92, 93, 117, 118, 255], dtype=np.uint8)
    out = FO.display_stretch(q, 105.0, 5.0)
    assert out[0] == 0 and out[3] == 255 and out[4] == 255
    # zero width band: inv = 0 -> every pixel maps to 0
    assert (FO.display_stretch(np.full(4, 9, np.uint8), 9.0, 0.0) == 0).all()


def test_display_scale_known_sizes():
    # max_h = 1024 - 320 = 704
    assert FO.display_scale(4096) == 6.0      # ceil(4096 / 704)
    assert FO.display_scale(1024) == 2.0
    assert FO.display_scale(512) == 1.0       # floor(704 / 512) = 1
    assert FO.display_scale(256) == 0.5       # floor(704 / 256) = 2
    assert FO.display_scale(64) == 1.0 / 11.0


def test_upscale_downscale_positions():
    a = np.arange(36, dtype=np.uint8).reshape(6, 6)
    d = FO.upscale_downscale(a, 4.0)
    assert d.tolist() == [[0, 4], [24, 28]]    # ceil(6 / 4) = 2 a side
    u = FO.upscale_downscale(a[:2, :2], 0.5)
    assert u.tolist() == [[0, 0, 1, 1], [0, 0, 1, 1], [6, 6, 7, 7], [6, 6, 7, 7]]
    assert FO.upscale_downscale(a, 1.0) is not None and FO.upscale_downscale(a, 1.0).shape == (6, 6)


def test_controller_known_trajectory():
    N, K0 = 512, 10.0
    c = FO.ViscController(N, K0)
    nu = 1e-4
    # first frame seeds both filters: p = 30, e = 0.5, de = 0.5, u = 0.125 + 12.5 -> clamp 0.01
    nu1, re1 = c.adapt(30.0 / 10000, nu)
    assert c._palin_filt == 30.0 and c._palin_filt_1 == 30.0 and c._e_prev == 0.5
    assert nu1 == pytest.approx(nu * math.exp(0.01), rel=1e-15)
    assert re1 == pytest.approx(1.0 / (nu * math.exp(0.01)), rel=1e-15)
    # second frame p_raw 36: filt1 = 30 + 6/6 = 31, filt = 30 + 1/6
    nu2, _ = c.adapt(36.0 / 10000, nu1)
    assert c._palin_filt_1 == pytest.approx(31.0, rel=1e-15)
    assert c._palin_filt == pytest.approx(30.0 + 1.0 / 6.0, rel=1e-15)
    e2 = (30.0 + 1.0 / 6.0 - 20.0) / 20.0
    assert c._e_prev == pytest.approx(e2, rel=1e-15)
    # u = 0.25 e2 + 25 (e2 - 0.5) = 0.127 + 0.208 -> clamp 0.01 again
    assert nu2 == pytest.approx(nu1 * math.exp(0.01), rel=1e-14)


def test_controller_dead_band_and_negative_clamp():
    c = FO.ViscController(512, 10.0)
    assert c.adapt(20.01 / 10000, 1e-4) is None        # |e| = 5e-4 < 1e-3
    assert c._e_prev == 0.0                            # the error is not recorded in the band
    c.reset()
    nu, _ = c.adapt(10.0 / 10000, 1e-4)                # e = -0.5 -> u clamps at -0.01
    assert nu == pytest.approx(1e-4 * math.exp(-0.01), rel=1e-15)


def test_controller_re_bounds():
    c = FO.ViscController(512, 10.0)
    re0 = float(FO.re_from_n_k0(512, 10.0))
    # a viscosity far below the band: Re clamps at 1000 Re0
    nu, re = c.adapt(30.0 / 10000, 1e-12)
    assert re == pytest.approx(1000.0 * re0, rel=1e-15) and nu == pytest.approx(1.0 / re, rel=1e-15)
    assert re0 == pytest.approx(10.0 ** (1.066240 * math.log10(512) - 0.083911 + 1.286396), rel=1e-14)
