"""Display frames built on the device: the per-frame work of the reference GUI.

For every rendered frame the reference GUI (turbo_main.py:1882-1935
``_on_timer``, 2075-2125 ``_update_image``) copies the full-resolution uint8
plane to the host (``get_frame_pixels``, turbo_wrapper.py:480-515) and then,
on the host:
  * takes the mean and standard deviation of those pixels and stops the run
    when sigma < 1 (turbo_main.py:2080-2087);
  * reads pal_over_ens_kmax2 (turbo_main.py:2015-2073, a device reduction and
    another transfer) and runs the viscosity controller
    (``_palin_filter_2nd`` / ``adapt_visc``, turbo_main.py:2154-2214);
  * builds the CSV row (``get_csv_tuple``, turbo_main.py:1965-1985: more
    reductions and transfers for flow_eddy_metrics and energy_joules);
  * stretches the pixels to [mu - 2.5 sigma, mu + 2.5 sigma] and scales them
    for display (turbo_main.py:2109-2114, 830-890).
``FrameLoop`` does all of that with one ``palx_frame`` call (kernels on the
handle stream, no host synchronisation inside) and one device-to-host copy
of the display-sized result: a record of scalars followed by the pixels.
The controller's new viscosity is written into the device scalars that the
next step reads, so stepping never waits for the host.
"""
from __future__ import annotations

import datetime as _dt
import math
import time
from dataclasses import dataclass

import numpy as np

from . import _palx as P
from .ensemble import re_from_n_k0

DISPLAY_NORM_K_STD = 2.5  # turbo_main.py:391
CSV_HEADER = (            # turbo_main.py:442-446
    "N", "K0", "Re", "CFL", "VISC", "STEPS", "PALIN", "SIG",
    "TIME", "DT", "MINUTES", "FPS", "U", "L", "TAU_L",
    "T_OVER_TAU_L", "E(J)", "TS",
)


def display_scale(N: int, screen_h: int = 1024, ui_margin: int = 320) -> float:
    """Integer display scale of turbo_main.py:830-856: > 1 downscales, < 1 upscales."""
    max_h = max(128, screen_h - ui_margin)
    if N >= max_h:
        return float(int(math.ceil(N / max_h)))
    up = max(1, int(math.floor(max_h / N)))
    return 1.0 / float(up)


def scale_arg(scale: float) -> int:
    """The palx_frame ``scale`` argument for a display scale (turbo_main.py:875-890)."""
    if scale == 1.0:
        return 1
    if scale < 1.0:
        return -int(round(1.0 / scale))
    return int(scale)


def frame_ctl(N: int, K0: float, target: float = 20.0, k_std: float = DISPLAY_NORM_K_STD) -> P.FrameCtl:
    """Controller constants of adapt_visc (turbo_main.py:2168-2214) for grid N, peak K0."""
    return P.FrameCtl(k_std=float(k_std), target=float(target), kp=0.25, ki=0.0, kd=25.0,
                      deadband=0.001, max_frac=0.01, palin_tau=5.0,
                      re0=float(re_from_n_k0(int(N), float(K0))), re_factor=1000.0)


@dataclass
class Frame:
    """One rendered frame: display pixels and the record of palx_frame."""
    pixels: np.ndarray  # (height, width) uint8
    record: np.ndarray  # float64[FRAME_NREC]

    def __getitem__(self, idx):
        return float(self.record[idx])

    @property
    def stopped(self) -> bool:
        return self.record[P.FR_STOP] != 0.0

    @property
    def mu(self) -> float:
        return float(self.record[P.FR_MU])

    @property
    def sig(self) -> float:
        return float(self.record[P.FR_SIGMA])

    @property
    def palinstrophy_over_enstrophy_kmax2(self) -> float:
        return float(self.record[P.FR_RATIO])


class FrameBuffer:
    """Device output buffer of palx_frame plus its pinned host copy."""

    def __init__(self, state, scale: int):
        import torch
        self.state = state
        self.scale = int(scale)
        self.width, self.height, self.nbytes = state.handle.frame_dims(self.scale)
        self.dev = torch.empty(self.nbytes, dtype=torch.uint8, device="cuda")
        self.host = torch.empty(self.nbytes, dtype=torch.uint8, pin_memory=True)
        self._npix = self.width * self.height

    def run(self, var: int, flags: int, ctl) -> Frame:
        """Enqueue one frame, copy record + pixels to the host once, and wait."""
        import torch
        S = self.state
        S.handle.frame(var, self.scale, flags, ctl, self.dev.data_ptr(), self.nbytes)
        with torch.cuda.stream(S.stream):
            self.host.copy_(self.dev, non_blocking=True)
        S.stream.synchronize()
        raw = self.host.numpy()
        nrec = P.FRAME_NREC * 8
        rec = raw[:nrec].view(np.float64).copy()
        pix = raw[nrec:nrec + self._npix].reshape(self.height, self.width).copy()
        return Frame(pix, rec)


def apply_record(sim, rec: np.ndarray, adapted_re: bool) -> None:
    """Mirror the device scalars of a frame record into the host state
    (sync_time_scalars_from_device, turbo_simulator.py:846-854; the Re / visc
    assignment of adapt_visc, turbo_main.py:2211-2214)."""
    S = sim.state
    S._dt = float(rec[P.FR_DT])
    S._cn = float(rec[P.FR_CN])
    S._visc = float(rec[P.FR_VISC])
    S._ur_valid = True
    sim.dt, sim.cn = S._dt, S._cn
    if adapted_re:
        sim.re = S.Re = float(rec[P.FR_RE])


class FrameLoop:
    """Headless counterpart of the reference GUI's timer loop.

    ``tick()`` is one timer tick (one DNS step, ``_on_timer``); every
    ``update_interval`` ticks it renders a frame.  ``render()`` can also be
    called directly.  ``adapt`` switches the viscosity controller on
    (the GUI's "visc adapt" button).
    """

    def __init__(self, sim, update_interval: int = 10, adapt: bool = False, target: float = 20.0,
                 stretch: bool = True, screen_h: int = 1024, ui_margin: int = 320,
                 k_std: float = DISPLAY_NORM_K_STD, spectrum_average=None):
        self.sim = sim
        self.update_interval = int(update_interval)
        self.adapt = bool(adapt)
        self.stretch = bool(stretch)
        self._target, self._k_std, self._screen = float(target), float(k_std), (screen_h, ui_margin)
        self.buf = None
        self._bind()
        self.spectrum_average = spectrum_average
        self.csv_rows: list[tuple] = []
        self.palinstrophy_over_enstrophy_kmax2 = None
        self.sig = None
        self.stopped = False
        self._status_update_counter = 0
        self._sim_start_time = time.time()
        self._sim_start_iter = sim.get_iteration()
        self.reset_controller()

    def _bind(self) -> None:
        """(Re)build the output buffer and constants for the simulator's current state,
        which set_N / reset_field replace (a new state starts a clean controller)."""
        sim = self.sim
        self.ctl = frame_ctl(sim.N, sim.k0, target=self._target, k_std=self._k_std)
        self.scale = scale_arg(display_scale(sim.N, *self._screen))
        self.buf = FrameBuffer(sim.state, self.scale)

    def reset_controller(self) -> None:
        """Clear the filters and errors (on_visc_adapt_toggled, turbo_main.py:1698-1705)."""
        self.sim.state.handle.frame_reset(float(self.sim.re))

    def render(self, var: int | None = None) -> Frame:
        sim = self.sim
        if self.buf.state is not sim.state:
            self._bind()
            self.reset_controller()
        var = sim.current_var if var is None else int(var)
        flags = 0
        if sim._next_dt_pending:
            flags |= P.FRAME_NEXTDT
        if self.stretch:
            flags |= P.FRAME_STRETCH
        if self.adapt:
            flags |= P.FRAME_ADAPT
        if self.spectrum_average is not None:
            from . import turbo_simulator as ts
            ts.sample_and_update_spectrum_average(sim.state, self.spectrum_average,
                                                  sim_time=float(sim.get_time()),
                                                  step=int(sim.get_iteration()), on_device=True)
        fr = self.buf.run(var, flags, self.ctl)
        sim._next_dt_pending = False
        apply_record(sim, fr.record, adapted_re=fr.record[P.FR_ADAPTED] != 0.0)
        self.sig = fr.sig
        if fr.stopped:  # turbo_main.py:2085-2087: the run stops, nothing else is updated
            self.stopped = True
            return fr
        self.palinstrophy_over_enstrophy_kmax2 = fr.palinstrophy_over_enstrophy_kmax2
        self.csv_rows.append(self.csv_tuple(fr))
        return fr

    def csv_tuple(self, fr: Frame) -> tuple:
        """get_csv_tuple (turbo_main.py:1965-1985) from a frame record."""
        sim = self.sim
        elapsed = time.time() - self._sim_start_time
        steps = sim.get_iteration() - self._sim_start_iter
        tau = fr[P.FR_TAU_L]
        t = float(sim.state.t)
        return (sim.N, int(sim.k0), float(sim.state.Re), sim.cfl, float(sim.state.visc),
                sim.get_iteration(), int(10000 * fr[P.FR_RATIO]), int(fr.sig), t, float(sim.state.dt),
                elapsed / 60.0, steps / elapsed if elapsed > 0 else float("inf"),
                fr[P.FR_U], fr[P.FR_L], tau, t / tau if tau > 0.0 else float("nan"), fr[P.FR_EJ],
                _dt.datetime.now(_dt.timezone.utc).strftime("%Y-%m-%dT%H:%M:%SZ"))

    def tick(self):
        """One timer tick (turbo_main.py:1882-1918): one step, a frame when due."""
        self.sim.step(self.update_interval)
        self._status_update_counter += 1
        if self._status_update_counter >= self.update_interval:
            self._status_update_counter = 0
            return self.render()
        return None
