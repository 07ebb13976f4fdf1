"""NumPy restatement of the reference GUI's per-frame work (test oracle only).

Each function follows the cited reference lines of ``palinstrophy/turbo_wrapper.py``
(``tw.py``) or ``palinstrophy/turbo_main.py`` (``tm.py``), in the same
arithmetic: NumPy float32 (or the plane's own dtype) with every operation
rounded on its own.  Pinned by the known answers in
``tests/test_frame_oracle.py`` (hand-derived values that a truncation /
rounding / scaling / filter mistake would change).

The one deliberate difference from the reference: ``pixel_stats`` returns
the exact mean and population standard deviation (float64 of exact integer
sums), which is what the device computes.  The reference takes them with
NumPy's float32 pairwise sums (tm.py:2082-2083), about 1e-7 relative away.
"""
from __future__ import annotations

import math

import numpy as np

DISPLAY_NORM_K_STD = 2.5  # tm.py:391


def float_to_pixels(field: np.ndarray) -> np.ndarray:
    """uint8 [1,255] map of a plane, GPU form (tw.py:317-337), in field.dtype."""
    f = np.asarray(field)
    dt = f.dtype.type
    fmin = f.min()
    fmax = f.max()
    rng = dt(fmax - fmin)
    is_const = abs(rng) <= np.float32(1.0e-12)
    denom = dt(1.0) if is_const else rng
    norm = ((f - fmin) / denom).astype(dt)
    pixf = (dt(1.0) + (norm * dt(254.0)).astype(dt)).astype(dt)
    pixf = np.clip(pixf, dt(1.0), dt(255.0))
    pix = pixf.astype(np.uint8)
    if is_const:
        pix = np.full(f.shape, 128, dtype=np.uint8)
    return pix


def pixel_stats(pix: np.ndarray):
    """Exact mean and population std of uint8 pixels (tm.py:2082-2083, see header)."""
    p = np.asarray(pix, dtype=np.int64)
    n = p.size
    s1 = int(p.sum())
    s2 = int((p * p).sum())
    mu = s1 / n
    var = max(0.0, (s2 - s1 * mu) / n)
    return mu, math.sqrt(var)


def display_stretch(pix: np.ndarray, mu: float, sig: float, k: float = DISPLAY_NORM_K_STD) -> np.ndarray:
    """[mu - k sig, mu + k sig] -> [0, 255] (tm.py:2109-2113), NumPy float32 semantics."""
    pix_f = np.asarray(pix).astype(np.float32, copy=False)
    lo = mu - k * sig
    hi = mu + k * sig
    inv = 255.0 / (hi - lo) if (hi - lo) != 0.0 else 0.0
    return ((pix_f - np.float32(lo)) * np.float32(inv)).round().clip(0.0, 255.0).astype(np.uint8)


def display_scale(N: int, screen_h: int = 1024, ui_margin: int = 320) -> float:
    """tm.py:830-856."""
    max_h = max(128, screen_h - ui_margin)
    if N >= max_h:
        return float(int(math.ceil(N / max_h)))
    up = int(math.floor(max_h / N))
    if up < 1:
        up = 1
    return 1.0 / float(up)


def upscale_downscale(pix: np.ndarray, scale: float) -> np.ndarray:
    """tm.py:875-890: nearest repeats for scale < 1, striding for scale > 1."""
    if scale == 1.0:
        return np.ascontiguousarray(pix)
    if scale < 1.0:
        up = int(round(1.0 / scale))
        return np.ascontiguousarray(np.repeat(np.repeat(pix, up, axis=0), up, axis=1))
    s = int(scale)
    return np.ascontiguousarray(pix[::s, ::s])


def re_from_n_k0(N, K0):
    """tm.py:2323-2328."""
    return 10.0 ** (1.066240 * np.log10(N) + -0.083911 * np.log10(K0) + 1.286396)


class ViscController:
    """_palin_filter_2nd and adapt_visc (tm.py:2154-2214) as a scalar state machine."""

    def __init__(self, N, K0, target=20.0):
        self.N, self.K0 = N, K0
        self._target = float(target)
        self.reset()

    def reset(self):
        """tm.py:1702-1705."""
        self._e_int = 0.0
        self._e_prev = 0.0
        self._palin_filt = 0.0
        self._palin_filt_1 = 0.0

    def _palin_filter_2nd(self, p_raw: float) -> float:
        palin_tau = 5.0
        alpha = 1.0 / (palin_tau + 1.0)
        if self._palin_filt == 0.0:
            self._palin_filt_1 = float(p_raw)
            self._palin_filt = float(p_raw)
        else:
            self._palin_filt_1 = float(self._palin_filt_1) + alpha * (float(p_raw) - float(self._palin_filt_1))
            self._palin_filt = float(self._palin_filt) + alpha * (float(self._palin_filt_1) - float(self._palin_filt))
        return float(self._palin_filt)

    def adapt(self, ratio: float, visc: float):
        """One rendered frame: returns (visc, Re) after the update, or None when unchanged."""
        deadband, max_frac = 0.001, 0.01
        Kp, Ki, Kd = 0.25, 0.0, 25.0
        p = self._palin_filter_2nd(float(10000 * ratio))
        e = (p - self._target) / self._target
        if abs(e) < deadband:
            return None
        if Ki != 0.0:
            self._e_int += e
        de = e - self._e_prev
        self._e_prev = e
        u = Kp * e + Ki * self._e_int + Kd * de
        u = max(-max_frac, min(max_frac, u))
        nu_new = float(visc) * math.exp(u)
        Re0 = float(re_from_n_k0(self.N, self.K0))
        factor = 1000.0
        Re_eff = max(Re0 / factor, min(factor * Re0, float(1.0 / nu_new)))
        return 1.0 / Re_eff, Re_eff
