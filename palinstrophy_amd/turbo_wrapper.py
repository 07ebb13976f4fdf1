"""MI355X counterpart of ``palinstrophy/turbo_wrapper.py`` (DnsSimulator).

Same public surface as the reference GUI wrapper (turbo_wrapper.py:15-536):
``step``, ``set_N``, ``reset_field``, ``get_frame_pixels``,
``make_pixels_component``, ``set_variable``, ``get_time``, ``get_iteration``,
``diagnostics``, ``save_png``.  Device arrays are torch tensors; no CuPy.
"""
from __future__ import annotations

import math
import os
from pathlib import Path
from time import perf_counter
from typing import Literal, Union

import numpy as np

from . import _palx as _P
from . import turbo_simulator as dns_all


class DnsSimulator:
    VAR_U = 0
    VAR_V = 1
    VAR_ENERGY = 2
    VAR_OMEGA = 3
    VAR_STREAM = 4

    def __init__(self, n: int = 512, re: float = 1000.0, k0: float = 15.0, cfl: float = 2.0,
                 backend: Literal["cpu", "gpu", "auto"] = "auto",
                 start_spectrum: dns_all.SPECTRUM = "KM3",
                 method: dns_all.TimeStepper = "LS_IMEX_RK3", precision: str | None = None):
        self.N = int(n)
        self.m = 3 * self.N
        self.re = float(re)
        self.k0 = float(k0)
        self.cfl = float(cfl)
        self.start_spectrum = start_spectrum
        self.method = method
        self.precision = precision
        self.seed = dns_all.pao_seed_from_env(1 + (int.from_bytes(os.urandom(8), "little") % dns_all.PAO_SEED_MAX))
        self.backend = backend
        self.max_steps = 5000
        self.fft_workers = 1  # the reference's scipy.fft pool size (CPU backend only, unused here)
        self.current_var = self.VAR_U
        self._build(self.seed, skip_pao=False)

    def _build(self, seed, skip_pao):
        start = perf_counter()
        self.state = dns_all.create_dns_state(N=self.N, Re=self.re, K0=self.k0, CFL=self.cfl,
                                              backend=self.backend, seed=seed, skip_pao=skip_pao,
                                              start_spectrum=self.start_spectrum, precision=self.precision)
        self.nx = int(self.state.NZ_full)
        self.ny = int(self.state.NX_full)
        self.py, self.px = self.nx, self.ny
        if not skip_pao:
            dns_all.dns_step2a(self.state)
            CFLM = float(dns_all.compute_cflm(self.state))
            self.state.dt = float(self.state.cflnum) / (CFLM * math.pi)
        self.state.cn = 1.0
        self.state.cnm1 = 0.0
        self.t = float(self.state.t)
        self.dt = float(self.state.dt)
        self.cn = float(self.state.cn)
        self.iteration = 0
        self._next_dt_pending = True
        print(f" DNS initialization took {perf_counter() - start:.3f} seconds")

    def step(self, mod_next_dt: int) -> None:
        """One DNS step; NEXTDT is deferred to the next frame (turbo_wrapper.py:122-167)."""
        S = self.state
        dt_old = S.dt
        if self.method == "LS_IMEX_RK3":
            dns_all.dns_step_ls_imex_rk3(S)
        elif self.method == "CNAB2":
            dns_all.dns_step2b(S)
            dns_all.dns_step3(S)
            dns_all.dns_step2a(S)
        else:
            raise ValueError(f"unknown time stepper: {self.method}")
        if (self.iteration % mod_next_dt) == 0:
            self._next_dt_pending = True
        S.t += dt_old
        self.t = float(S.t)
        self.dt = float(S.dt)
        self.cn = float(S.cn)
        self.iteration += 1
        S.it = int(self.iteration)

    def set_N(self, N: int, skip_pao: bool = False) -> None:
        self.N = int(N)
        self.m = 3 * self.N
        self._build(self.seed, skip_pao)

    def reset_field(self) -> None:
        self.seed = dns_all.pao_seed_from_env(1 + (int.from_bytes(os.urandom(8), "little") % dns_all.PAO_SEED_MAX))
        self._build(self.seed, skip_pao=False)

    def diagnostics(self) -> dict:
        return {"t": float(self.t), "dt": float(self.dt), "cn": float(self.cn)}

    @staticmethod
    def _float_to_pixels_gpu(field):
        """uint8 [1,255] normalisation on the device (turbo_wrapper.py:317-337)."""
        import torch
        fmin = field.min()
        fmax = field.max()
        rng = fmax - fmin
        is_const = torch.abs(rng) <= 1.0e-12
        denom = torch.where(is_const, torch.ones_like(rng), rng)
        pix = torch.clamp(1.0 + (field - fmin) / denom * 254.0, 1.0, 255.0).to(torch.uint8)
        return torch.where(is_const, torch.full_like(pix, 128), pix)

    def _component(self, var):
        S = self.state
        if var == self.VAR_ENERGY:
            dns_all.dns_kinetic(S)
            return S.ur_full[2]
        if var == self.VAR_OMEGA:
            dns_all.dns_om2_phys(S)
            return S.ur_full[2]
        if var == self.VAR_STREAM:
            dns_all.dns_stream_func(S)
            return S.ur_full[2]
        return S.ur_full[1 if var == self.VAR_V else 0]

    def make_pixels(self, comp: int = 1) -> np.ndarray:
        idx = int(comp) - 1
        if idx < 0 or idx > 2:
            idx = 0
        return self._float_to_pixels_gpu(self.state.ur_full[idx]).cpu().numpy()

    def _raw_frame(self, var: int, nextdt: bool):
        """Full-resolution [1,255] pixels of var through palx_frame (one copy to the host)."""
        from .frame import FrameBuffer
        buf = getattr(self, "_frame_buf", None)
        if buf is None or buf.state is not self.state:
            buf = self._frame_buf = FrameBuffer(self.state, 1)
        return buf.run(int(var), _P.FRAME_NEXTDT if nextdt else 0, None)

    def make_pixels_component(self, var: int | None = None) -> np.ndarray:
        var = self.current_var if var is None else var
        fr = self._raw_frame(var, False)
        self.state._ur_valid = True
        return fr.pixels

    def get_frame_pixels(self) -> np.ndarray:
        """uint8 [1,255] plane of the current variable; runs the deferred NEXTDT
        in the same device pass (turbo_wrapper.py:480-515)."""
        from .frame import apply_record
        fr = self._raw_frame(self.current_var, self._next_dt_pending)
        self._next_dt_pending = False
        apply_record(self, fr.record, adapted_re=False)
        if math.isnan(fr.record[_P.FR_CFLM]):  # as sync_time_scalars_from_device
            raise FloatingPointError("NaN in the velocity field (CFL reduction)")
        return fr.pixels

    def set_variable(self, var: int) -> None:
        self.current_var = int(var)

    def get_time(self) -> float:
        return float(self.t)

    def get_iteration(self) -> int:
        return self.iteration

    def save_png(self, path: Union[str, Path], comp: int = 1) -> None:
        from PIL import Image
        Image.fromarray(self.make_pixels(comp), mode="L").save(str(path))
