"""MI355X-native counterpart of ``palinstrophy/turbo_simulator.py``.

Same module-level API as the reference (names, arguments, error types), so
callers written for the reference -- the ``sim`` CLI loop, the Qt wrapper --
run unchanged.  Every array operation of the DNS step runs in libpalx.so
(hand-written HIP for gfx950) through ctypes; this module only keeps host
scalars, the NEXTDT cadence and printing.  There is no CPU fallback:
``backend="cpu"`` raises, and a missing libpalx.so raises at state creation.

Differences from the reference, all deliberate:
  * ``precision`` selects FP32 (the reference's own arithmetic, default) or
    FP64 for the bulk FFT / nonlinear work; t, dt, CFL and diagnostic sums are
    FP64 on the device in both cases.
  * Array attributes (``om2``, ``fnm1``, ``ur_full``) are device arrays
    (``devarray.DeviceArray``: torch tensors with CuPy's ``.get()``,
    ``numpy.asarray`` and ``.size``) that view the device buffers without
    copying (live, like the reference's CuPy arrays; ``om2``/``fnm1`` are
    transposed views of column-major storage); assigning to the attribute
    copies into the buffers.  ``ur_full`` is materialised on access (the step
    itself never writes the padded physical planes).  ``xp`` is the
    CuPy-style namespace of ``devarray``.
  * ``uc`` (compact velocity spectra (N, 3N/4+1, 3)) is computed from the
    current omega on access, as the reference's would be right after a step;
    an assigned ``uc`` is returned until the next step.  ``uc_full``
    (3, 3N/2, 3N/4+1) holds the current velocity spectra when read; if it has
    been handed out since the last step, the next ``dns_step2a`` rebuilds the
    velocity from its rows z < N, kx < N/2, which is how the reference GUI
    loads a restart (turbo_main.py:1443-1445 then dns_step2a).  The step
    itself keeps the velocity in z-inverse-transformed work planes.
"""
from __future__ import annotations

import csv
import datetime as _dt
import io
import math
import os
import sys
import time
from dataclasses import dataclass, field
from typing import Literal, cast

import numpy as _np

from . import _palx as _P

SCIPYTURBO_SEED_ENV = "SCIPYTURBO_SEED"
PAO_SEED_MIN = 1
PAO_SEED_MAX = 5010
TimeStepper = Literal["CNAB2", "LS_IMEX_RK3", "CHECK"]
SPECTRUM = Literal["PAO", "KM3"]
PRECISION_ENV = "PALINSTROPHY_PRECISION"
TUNE_ENV = "PALINSTROPHY_TUNE"  # "1": autotune the step kernels of every new state
_METHOD_ID = {"CNAB2": _P.CNAB2, "LS_IMEX_RK3": _P.LS_IMEX_RK3}


def pao_seed_from_env(default: int | None = None) -> int | None:
    """Seed from SCIPYTURBO_SEED in [1, 5010] (turbo_simulator.py:54-73)."""
    raw = os.environ.get(SCIPYTURBO_SEED_ENV)
    if raw is None or raw.strip() == "":
        return default
    try:
        seed = int(raw.strip(), 10)
    except ValueError as exc:
        raise ValueError(
            f"{SCIPYTURBO_SEED_ENV} must be an integer from {PAO_SEED_MIN} to {PAO_SEED_MAX}; got {raw!r}"
        ) from exc
    if not PAO_SEED_MIN <= seed <= PAO_SEED_MAX:
        raise ValueError(f"{SCIPYTURBO_SEED_ENV} must be from {PAO_SEED_MIN} to {PAO_SEED_MAX}; got {seed}")
    return seed


def _torch():
    import torch
    return torch


def _xp():
    from . import devarray
    return devarray


def get_xp(backend: Literal["cpu", "gpu", "auto"] = "auto"):
    """Array module of the device state: torch (turbo_simulator.py:547-566)."""
    if backend == "cpu":
        raise RuntimeError("this implementation runs the DNS on the GPU only; use backend='gpu' or 'auto'")
    return _torch()


def _precision_bytes(precision) -> int:
    if precision is None:
        precision = os.environ.get(PRECISION_ENV, "fp32")
    p = str(precision).lower()
    if p in ("fp32", "float32", "f32", "4", "single"):
        return 4
    if p in ("fp64", "float64", "f64", "8", "double"):
        return 8
    raise ValueError(f"precision must be 'fp32' or 'fp64'; got {precision!r}")


class _DeviceBuffer:
    """__cuda_array_interface__ over libpalx device memory; holds the handle
    so the memory outlives every tensor viewing it."""

    def __init__(self, owner, ptr, shape, typestr, strides):
        self._owner = owner
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": typestr, "data": (int(ptr), False),
                                         "version": 3, "strides": strides}


class DnsState:
    """Device state of one DNS (turbo_simulator.py:654-810 DnsState)."""

    def __init__(self, N, Re, K0, CFL, seed, start_spectrum, precision, populate_compact_ur):
        self.backend = "gpu"
        self.xp = _xp()
        self.Nbase = int(N)
        self.NX = self.NZ = int(N)
        self.NK = 3 * self.NX // 4 + 1
        self.NX_full = self.NZ_full = 3 * self.NX // 2
        self.NK_full = self.NX_full // 2 + 1
        self.Re = float(Re)
        self.K0 = float(K0)
        self.start_spectrum = start_spectrum
        self.seed_init = int(seed)
        self.fft_workers = 1
        self.populate_compact_ur = bool(populate_compact_ur)
        self.inv_dx = float(N) / (2.0 * math.pi)
        self.precision = precision
        self.handle = _P.Handle(N, precision)
        torch = _torch()
        self.stream = torch.cuda.current_stream()  # the stream every call of the handle runs on
        self.handle.call("palx_set_stream", self.stream.cuda_stream)
        self._views = {}
        self.t = 0.0
        self.it = 0
        self.pao_stats = None
        self._dt = 0.0
        self._cn = 1.0
        self._cnm1 = 0.0
        self._visc = 0.0
        self._cflnum = float(CFL)
        self._ur_valid = False
        self.handle.set_scalar(_P.S_CFLNUM, self._cflnum)
        self.handle.set_scalar(_P.S_CN, 1.0)
        alfa = _np.arange(N // 2, dtype=_np.float32)
        gamma = _np.array([z if z < N // 2 else z - N for z in range(N)], dtype=_np.float32)
        self.alfa = self.xp.asarray(alfa)
        self.gamma = self.xp.asarray(gamma)
        self.step3_K2 = self.xp.wrap((self.alfa[None, :] ** 2 + self.gamma[:, None] ** 2).contiguous())
        self._uc_full = None          # allocated when first read
        self._uc_full_pending = False  # handed out since the last step: dns_step2a loads from it
        self._uc_assigned = None       # an assigned uc, returned until the next step

    # --- host-writable scalars mirror FP64 device scalars (ts.py:742-753)
    def _scalar_prop(idx, attr):  # noqa: N805
        def get(self):
            return getattr(self, attr)

        def set_(self, value):
            setattr(self, attr, float(value))
            self.handle.set_scalar(idx, float(value))
        return property(get, set_)

    dt = _scalar_prop(_P.S_DT, "_dt")
    cn = _scalar_prop(_P.S_CN, "_cn")
    cnm1 = _scalar_prop(_P.S_CNM1, "_cnm1")
    visc = _scalar_prop(_P.S_VISC, "_visc")
    cflnum = _scalar_prop(_P.S_CFLNUM, "_cflnum")
    del _scalar_prop

    # --- arrays: live zero-copy torch views of the device state, like the
    # reference's CuPy arrays; assignment copies into the device buffers.
    def _view(self, field_id):
        v = self._views.get(field_id)
        if v is None:
            ptr, _ = self.handle.device_ptr(field_id)
            N, H, M = self.Nbase, self.Nbase // 2, self.NX_full
            real = self.precision == 4
            isz = 4 if real else 8
            if field_id == _P.F_UR:
                shape, typestr, strides = (3, M, M), ("<f4" if real else "<f8"), None
            elif field_id == _P.F_UC0:  # compact velocity spectra [c][z][kx]
                shape, typestr, strides = (2, N, H), ("<c8" if real else "<c16"), None
            else:  # omega / fnm1 are stored column-major [kx][z]: a transposed (N, H) view
                shape, typestr, strides = (N, H), ("<c8" if real else "<c16"), (2 * isz, 2 * isz * N)
            t = _torch().as_tensor(_DeviceBuffer(self.handle, ptr, shape, typestr, strides), device="cuda")
            v = self.xp.wrap(t)
            self._views[field_id] = v
        return v

    def _set(self, field_id, value):
        if hasattr(value, "is_cuda") and value.is_cuda:
            self._view(field_id).copy_(value)  # device to device, on the handle stream
            return
        a = value.detach().cpu().numpy() if hasattr(value, "detach") else _np.asarray(value)
        self.handle.upload(field_id, a)

    @property
    def om2(self):
        return self._view(_P.F_OM)

    @om2.setter
    def om2(self, v):
        self._set(_P.F_OM, v)

    @property
    def fnm1(self):
        return self._view(_P.F_FNM1)

    @fnm1.setter
    def fnm1(self, v):
        self._set(_P.F_FNM1, v)

    @property
    def ur_full(self):
        if not self._ur_valid:
            self.handle.physical()
            self._ur_valid = True
        return self._view(_P.F_UR)

    # --- the reference's velocity spectra (turbo_simulator.py:944-947)
    def _velocity_from_omega(self):
        """u_hat, w_hat (N, N/2) from omega on the device (turbo_simulator.py:1555-1585)."""
        torch = _torch()
        om = self.om2
        rdt = torch.float32 if self.precision == 4 else torch.float64
        kx = self.alfa.to(rdt)[None, :]
        kz = self.gamma.to(rdt)[:, None]
        inv = 1.0 / (kx * kx + kz * kz + 1.0e-30)
        u = -1j * (kz * inv) * om
        w = 1j * (kx * inv) * om
        kz0 = self.gamma.to(rdt)
        inv_kz = torch.where(kz0 != 0, 1.0 / torch.where(kz0 != 0, kz0, torch.ones_like(kz0)), torch.zeros_like(kz0))
        u[:, 0] = -1j * inv_kz * om[:, 0]
        w[:, 0] = 0
        return u, w

    @property
    def uc(self):
        """Compact velocity spectra (N, 3N/4+1, 3), complex64, components u, w, 0."""
        if self._uc_assigned is not None:
            return self._uc_assigned
        torch = _torch()
        N, H = self.Nbase, self.Nbase // 2
        u, w = self._velocity_from_omega()
        out = torch.zeros((N, self.NK, 3), dtype=torch.complex64, device="cuda")
        out[:, :H, 0] = u
        out[:, :H, 1] = w
        return self.xp.wrap(out)

    @uc.setter
    def uc(self, value):
        self._uc_assigned = self.xp.asarray(value)

    @property
    def uc_full(self):
        """(3, 3N/2, 3N/4+1) spectra: rows z < N, kx < N/2 of components 0, 1
        hold the current velocity spectra on reading; writes take effect at
        the next dns_step2a (turbo_main.py:1443-1445)."""
        torch = _torch()
        M, N, H = self.NX_full, self.Nbase, self.Nbase // 2
        cdt = torch.complex64 if self.precision == 4 else torch.complex128
        if self._uc_full is None:
            self._uc_full = self.xp.wrap(torch.zeros((3, M, self.NK_full), dtype=cdt, device="cuda"))
        if not self._uc_full_pending:
            u, w = self._velocity_from_omega()
            self._uc_full.zero_()
            self._uc_full[0, :N, :H] = u
            self._uc_full[1, :N, :H] = w
            self._uc_full_pending = True
        return self._uc_full

    def _load_velocity_from_uc_full(self):
        """Velocity work planes from uc_full rows z < N, kx < N/2 (the dealias
        and row move of dns_step2a, turbo_simulator.py:2020-2034, keep exactly
        these), through the handle's velocity-spectra buffer."""
        N, H = self.Nbase, self.Nbase // 2
        self._view(_P.F_UC0).copy_(self._uc_full[0:2, :N, :H])
        self.handle.load_velocity()
        self._uc_full_pending = False

    def _stepped(self):
        """State after a step: the velocity lives in the work planes again."""
        self._ur_valid = False
        self._uc_full_pending = False
        self._uc_assigned = None

    def sync(self):
        self.handle.sync()

    def flow_eddy_metrics(self) -> tuple[float, float, float]:
        """U, L, tau_L (turbo_simulator.py:764-810), FP64 device sums."""
        d = self.handle.diag()
        self._ur_valid = True
        return float(d[0]), float(d[1]), float(d[2])

    def diagnostics(self) -> dict:
        d = self.handle.diag()
        self._ur_valid = True
        U, L, tau, Z, Pal, kmax2, Ej, E = (float(x) for x in d)
        pe = Pal / (Z * kmax2) if (Z > 0 and kmax2 > 0) else 0.0
        return {"U": U, "L": L, "tau_L": tau, "enstrophy": Z, "palinstrophy": Pal,
                "kmax2": kmax2, "pal_over_ens_kmax2": pe, "energy_joules": Ej, "energy_sum": E}


def create_dns_state(
    N: int = 8,
    Re: float = 1e5,
    K0: float = 100.0,
    CFL: float = 0.75,
    backend: Literal["cpu", "gpu", "auto"] = "auto",
    seed: int = 1,
    skip_pao: bool = False,
    populate_compact_ur: bool = True,
    start_spectrum: SPECTRUM = "KM3",
    precision: str | None = None,
) -> DnsState:
    """Allocate the device state and seed it (turbo_simulator.py:878-1054)."""
    get_xp(backend)
    prec = _precision_bytes(precision)
    S = DnsState(N, Re, K0, CFL, seed, start_spectrum, prec, populate_compact_ur)
    if os.environ.get(TUNE_ENV, "0") not in ("", "0"):
        # kernel variants from the autotuner (cached per device, library and
        # size); without it the handle keeps its measured defaults
        from . import autotune
        autotune.tune_handle(S.handle, reps=5)
    dns_pao_host_init(S, skip_pao=skip_pao)
    S.dt = 0.0
    S.cn = 1.0
    S.cnm1 = 0.0
    return S


def dns_pao_host_init(S: DnsState, skip_pao: bool = False):
    """Seeded PAO/KM3 field -> omega_0 and work planes (turbo_simulator.py:1059-1239)."""
    N = S.Nbase
    K0 = _np.float32(S.K0)
    km3 = S.start_spectrum == "KM3"
    label = "KM3 k^-3" if km3 else "PAO k*exp(-(k/K0)^2)"
    print("--- INITIALIZING MI355X/HIP ---", _dt.datetime.now().strftime("%Y-%m-%d %H:%M"))
    print(f" N={N}, K0={int(K0)}, Re={S.Re:,.1f}")
    print(f" Start spec. = {label}")
    if skip_pao:
        S.visc = float(_np.float32(1.0 / float(_np.float32(S.Re))))
        print(" PAO spectrum skipped (loading saved case)")
        return
    uc0, st = _P.pao_generate(N, float(S.K0), float(S.Re), S.seed_init, km3)
    om0 = _P.calcom_f32(uc0)
    S.pao_stats = st
    S.visc = float(st[4])
    Q2, W2 = float(st[1]), float(st[2])
    A1, A3, A4 = float(st[5]), float(st[7]), float(st[8])
    visc = S.visc
    EP = visc * W2
    KOL = (visc ** 3 / EP) ** 0.25 if EP > 0 else float("nan")
    DXZ = float(_np.float32(2.0) * _np.float32(3.14159265358979) / _np.float32(N))
    print(f" N           = {N:.8g}")
    print(f" Reynolds n. = {float(S.Re):.8g}")
    print(f" K0          = {K0:.8g}")
    print(f" Energy      = {Q2:.8g}")
    print(f" WiWi        = {W2:.8g}")
    print(f" Time scale  = {0.5 * Q2 / EP:.8g}")
    print(f" Kolmogorov  = {KOL:.8g}")
    print(f" Viscosity   = {visc:.8g}")
    print(f" dx/Kol.     = {DXZ / KOL:.8g}")
    print(f" PAO seed in = {S.seed_init:.8g}")
    print(f" PAO seed out= {int(st[0]):.8g}")
    S.handle.load_initial(uc0, om0)
    S._ur_valid = False
    print(f" {S.start_spectrum} INITIALIZATION OK. VISC={visc:.8g}")


def dns_load_spectral(S: DnsState, om2, fnm1=None):
    """Load omega (and fnm1) and rebuild u_hat, w_hat from omega.

    Restart entry point: the reference restores the *initial* compact
    velocity spectra (turbo_main.py:1413-1420, stale after the first step);
    here the velocity is always reconstructed from the loaded omega.
    """
    S.om2 = om2
    if fnm1 is not None:
        S.fnm1 = fnm1
    # the reconstruction of the step's column pass (turbo_simulator.py:1555-1585)
    S.handle.recon_velocity()
    S._stepped()


def dns_step2b(S: DnsState) -> None:
    """Products u^2-w^2, uw and their x-forward (turbo_simulator.py:1378-1513)."""
    S.handle.stage_products()
    S._stepped()


def dns_step3(S: DnsState, fuse: bool = True) -> None:
    """CNAB2 update + velocity reconstruction + padded z-inverse; cnm1 = cn
    afterwards (turbo_simulator.py:1588-1771)."""
    S.handle.stage_update(_P.CNAB2, 0)
    S._cnm1 = S._cn
    S._stepped()


def dns_step_ls_imex_rk3(S: DnsState) -> None:
    """One LS-IMEX-RK3 step, three fused stages (turbo_simulator.py:1900-1943).

    The stages run as one step (STAGE_IN_STEP): stages 0 and 1 hand the next
    stage its pre-update term through fnm1's storage instead of storing
    omega and fn, which no caller of the reference sees between stages
    either; omega, fnm1 and the velocity afterwards are those of three plain
    stages, bit for bit."""
    for stage in range(3):
        S.handle.stage_products()
        S.handle.stage_update(_P.LS_IMEX_RK3, stage | _P.STAGE_IN_STEP)
    S._stepped()


def dns_step2a(S: DnsState) -> None:
    """Inverse to physical space (turbo_simulator.py:1949-2100).

    The z half runs inside dns_step3 / the RK3 stage; the x half is deferred
    until ``ur_full`` or the CFL is needed, so this only invalidates the
    cached physical planes -- unless ``uc_full`` was handed out since the last
    step: then the velocity is rebuilt from it (restart load,
    turbo_main.py:1443-1450)."""
    if S._uc_full_pending:
        S._load_velocity_from_uc_full()
    S._ur_valid = False


def compute_cflm(S: DnsState):
    """max(|u|+|w|) * N/2pi over the padded grid (turbo_simulator.py:2106-2125)."""
    c = S.handle.cflm()
    S._ur_valid = True
    return c


def next_dt(S: DnsState, sync_host: bool = False):
    """CFL-adaptive dt on the device (turbo_simulator.py:2127-2167)."""
    S.handle.next_dt()
    S._ur_valid = True
    if sync_host:
        sync_time_scalars_from_device(S)
        return float(S.handle.scalars()[_P.S_CFLM])
    return None


def sync_time_scalars_from_device(S: DnsState) -> None:
    """Refresh host dt/cn/cnm1 (turbo_simulator.py:846-854).

    Raises FloatingPointError when the device CFL reduction has met a NaN
    (a blown-up field); the reference has no such check and keeps stepping."""
    sc = S.handle.scalars()
    S._dt, S._cn, S._cnm1 = float(sc[_P.S_DT]), float(sc[_P.S_CN]), float(sc[_P.S_CNM1])
    if sc[_P.S_NAN] != 0.0:
        raise FloatingPointError("NaN in the velocity field (CFL reduction)")


# ---------------------------------------------------------------- diagnostics / spectrum
@dataclass
class SpectrumSample:
    nbins: int = 0
    r_max: float = 0.0
    k_nyq: float = 0.0
    energy: _np.ndarray = field(default_factory=lambda: _np.zeros(0, dtype=_np.float64))
    count: _np.ndarray = field(default_factory=lambda: _np.zeros(0, dtype=_np.int32))


@dataclass
class SpectrumAverageState:
    initialized: bool = False
    has_previous: bool = False
    since_restart: bool = False
    Nbase: int = 0
    NX_full: int = 0
    NZ_full: int = 0
    NK_full: int = 0
    nbins: int = 0
    sample_count: int = 0
    first_step: int = 0
    last_step: int = 0
    r_max: float = 0.0
    k_nyq: float = 0.0
    total_dt: float = 0.0
    previous_time: float = 0.0
    first_time: float = 0.0
    last_time: float = 0.0
    previous_energy: _np.ndarray = field(default_factory=lambda: _np.zeros(0, dtype=_np.float64))
    integral_energy: _np.ndarray = field(default_factory=lambda: _np.zeros(0, dtype=_np.float64))
    count: _np.ndarray = field(default_factory=lambda: _np.zeros(0, dtype=_np.int32))


def compute_energy_spectrum_uv_bins(S: DnsState, on_device: bool = False) -> SpectrumSample:
    """Radial energy spectrum from omega (turbo_simulator.py:2360-2383).

    on_device: energy / count are float64 / int32 torch tensors filled on the
    device (no host transfer, no synchronisation); else numpy arrays."""
    if S is None:
        raise ValueError("DnsState is required")
    n = int(S.Nbase)
    nbins = max(32, 2 * n)
    if on_device:
        torch = _torch()
        e = torch.empty(nbins, dtype=torch.float64, device="cuda")
        c = torch.empty(nbins, dtype=torch.int32, device="cuda")
        S.handle.spectrum_into(e.data_ptr(), c.data_ptr(), nbins)
    else:
        e, c = S.handle.spectrum(nbins)
    return SpectrumSample(nbins=nbins, r_max=math.sqrt(2.0), k_nyq=0.5 * float(n), energy=e, count=c)


# The average below runs on numpy arrays (host samples, as the reference) or on
# torch CUDA tensors (device samples: the trapezoid sum stays on the device and
# only the read-out in spectrum_average_mean_energy / the CSV writer copies).
def _is_tensor(x) -> bool:
    return hasattr(x, "is_cuda")


def _copy(x):
    return x.clone() if _is_tensor(x) else _np.array(x, copy=True)


def _as_f64(x):
    return x.double() if _is_tensor(x) else _np.asarray(x, dtype=_np.float64)


def _numel(x) -> int:
    return int(x.numel()) if _is_tensor(x) else int(_np.asarray(x).size)


def _as_i32(x):
    if _is_tensor(x):
        import torch
        return x.to(torch.int32).clone()
    return _np.asarray(x, dtype=_np.int32).copy()


def _host(x):
    return x.detach().cpu().numpy() if _is_tensor(x) else _np.asarray(x)


def mark_spectrum_average_since_restart(average: SpectrumAverageState, S: DnsState | None) -> None:
    average.__dict__.update(SpectrumAverageState().__dict__)
    average.since_restart = True
    if S is not None:
        average.Nbase, average.NX_full = int(S.Nbase), int(S.NX_full)
        average.NZ_full, average.NK_full = int(S.NZ_full), int(S.NK_full)


def _geometry_matches(average, S, sample) -> bool:
    return (average.Nbase == int(S.Nbase) and average.NX_full == int(S.NX_full)
            and average.NZ_full == int(S.NZ_full) and average.NK_full == int(S.NK_full)
            and average.nbins == int(sample.nbins)
            and tuple(average.previous_energy.shape) == tuple(sample.energy.shape)
            and tuple(average.integral_energy.shape) == tuple(sample.energy.shape)
            and tuple(average.count.shape) == tuple(sample.count.shape))


def initialize_spectrum_average_from_sample(average, S, sample, sim_time, step) -> None:
    keep = bool(average.since_restart)
    average.initialized = True
    average.has_previous = True
    average.since_restart = keep
    average.Nbase, average.NX_full = int(S.Nbase), int(S.NX_full)
    average.NZ_full, average.NK_full = int(S.NZ_full), int(S.NK_full)
    average.nbins = int(sample.nbins)
    average.sample_count = 1
    average.first_step = average.last_step = int(step)
    average.r_max, average.k_nyq = float(sample.r_max), float(sample.k_nyq)
    average.total_dt = 0.0
    average.previous_time = average.first_time = average.last_time = float(sim_time)
    average.previous_energy = _copy(_as_f64(sample.energy))
    average.integral_energy = average.previous_energy * 0.0
    average.count = _as_i32(sample.count)


def update_spectrum_average(average, S, sample, sim_time, step) -> bool:
    """Trapezoid-in-time accumulation of spectrum samples (turbo_simulator.py:2443-2497)."""
    if sample.nbins <= 0 or _numel(sample.energy) == 0 or tuple(sample.energy.shape) != tuple(sample.count.shape):
        return False
    if not average.initialized:
        initialize_spectrum_average_from_sample(average, S, sample, sim_time, step)
        return True
    if not _geometry_matches(average, S, sample):
        return False
    energy = _as_f64(sample.energy)
    if _is_tensor(energy) != _is_tensor(average.previous_energy):
        return False  # host and device samples do not mix in one average
    average.count = _as_i32(sample.count)
    average.r_max, average.k_nyq = float(sample.r_max), float(sample.k_nyq)
    sim_time, step = float(sim_time), int(step)
    if not average.has_previous:
        average.has_previous = True
        average.previous_energy = _copy(energy)
        average.previous_time = average.first_time = average.last_time = sim_time
        average.first_step = average.last_step = step
        average.sample_count = 1
        return True
    if sim_time <= average.previous_time or step == average.last_step:
        average.previous_energy = _copy(energy)
        average.previous_time = average.last_time = sim_time
        average.last_step = step
        if average.sample_count == 0:
            average.sample_count = 1
            average.first_time = sim_time
            average.first_step = step
        return True
    dt = sim_time - average.previous_time
    average.integral_energy += 0.5 * dt * (average.previous_energy + energy)
    average.total_dt += dt
    average.previous_energy = _copy(energy)
    average.previous_time = average.last_time = sim_time
    average.last_step = step
    average.sample_count += 1
    return True


def sample_and_update_spectrum_average(S, average, sim_time=None, step=None, on_device=None) -> bool:
    """on_device None: keep where the average already lives (host for a new one,
    as the reference); True: sample and accumulate on the device."""
    if on_device is None:
        on_device = _is_tensor(average.previous_energy)
    sample = compute_energy_spectrum_uv_bins(S, on_device=on_device)
    return update_spectrum_average(average, S, sample, float(S.t if sim_time is None else sim_time),
                                   int(S.it if step is None else step))


def spectrum_average_kind(average: SpectrumAverageState) -> str:
    if average.total_dt > 0.0:
        return ("simulation-time trapezoid since restart" if average.since_restart
                else "simulation-time trapezoid")
    return "single snapshot since restart" if average.since_restart else "single snapshot"


def spectrum_average_mean_energy(average: SpectrumAverageState) -> _np.ndarray:
    if not average.initialized or not average.has_previous or _host(average.previous_energy).size == 0:
        raise ValueError("no spectrum samples accumulated")
    if average.total_dt > 0.0:
        return _host(average.integral_energy) / average.total_dt
    return _host(average.previous_energy).copy()


def _csv_comment_line(values) -> str:
    buf = io.StringIO()
    csv.writer(buf, lineterminator="").writerow(values)
    return "# " + buf.getvalue() + "\n"


def save_energy_spectrum_csv_data(filename, energy, count, r_max, meta_header=None, meta_row=None,
                                  average=None, average_kind=None) -> None:
    """energy_spectrum.csv writer (turbo_simulator.py:2532-2585)."""
    e = _np.asarray(energy, dtype=_np.float64)
    c = _np.asarray(count, dtype=_np.int64)
    nbins = min(int(e.size), int(c.size))
    header = list(meta_header or ())
    values = list(meta_row or ())
    if average is not None and average_kind is not None:
        header += ["AVERAGE_KIND", "SPECTRUM_SAMPLES", "T_START", "T_END", "TOTAL_DT", "FIRST_STEP", "LAST_STEP"]
        values += [average_kind, int(average.sample_count), float(average.first_time),
                   float(average.last_time), float(average.total_dt), int(average.first_step),
                   int(average.last_step)]
    with open(filename, "w", newline="") as f:
        if header and values:
            f.write(_csv_comment_line(header))
            f.write(_csv_comment_line(values))
        w = csv.writer(f)
        w.writerow(("normalized_radius", "shell_sum_energy", "count"))
        for i in range(nbins):
            if c[i] > 0 and e[i] > 0.0:
                r0 = float(r_max) * float(i) / float(nbins)
                r1 = float(r_max) * float(i + 1) / float(nbins)
                w.writerow((0.5 * (r0 + r1), e[i], int(c[i])))
    label = f"{average_kind} energy spectrum" if average_kind else "energy spectrum"
    print(f"[CSV] Wrote {filename} ({label}, {nbins} bins)")


def save_energy_spectrum_uv_csv(S, filename, meta_header=None, meta_row=None) -> None:
    s = compute_energy_spectrum_uv_bins(S)
    save_energy_spectrum_csv_data(filename, s.energy, s.count, s.r_max, meta_header, meta_row)


def save_spectrum_average_csv(average, filename, meta_header=None, meta_row=None) -> None:
    save_energy_spectrum_csv_data(filename, spectrum_average_mean_energy(average), _host(average.count),
                                  average.r_max, meta_header, meta_row, average,
                                  spectrum_average_kind(average))


# ---------------------------------------------------------------- field getters
def dns_kinetic(S: DnsState) -> None:
    """ur_full[2] = sqrt(u^2 + w^2) (turbo_simulator.py:2629-2636)."""
    S.handle.getter(_P.VAR_ENERGY)
    S._ur_valid = True


def dns_om2_phys(S: DnsState) -> None:
    """ur_full[2] = physical vorticity (turbo_simulator.py:2684-2687)."""
    S.handle.getter(_P.VAR_OMEGA)
    S._ur_valid = True


def dns_stream_func(S: DnsState) -> None:
    """ur_full[2] = physical stream function (turbo_simulator.py:2690-2708)."""
    S.handle.getter(_P.VAR_STREAM)
    S._ur_valid = True


def vfft_full_inverse_uc_full_to_ur_full(S: DnsState) -> None:
    """u, w on the padded grid into ur_full[0:2] (turbo_simulator.py:1246-1273).

    The velocity spectra live in the work planes between steps, so this is
    the x half of the padded inverse (the physical pass)."""
    S.handle.physical()
    S._ur_valid = True


def dump_field_as_pgm_full(S: DnsState, comp: int, filename: str) -> None:
    """Binary PGM of ur_full[comp] (turbo_simulator.py:2173-2222): the plane as
    float32, mapped to [1,255] in double with round half up, constant -> 128."""
    field = S.ur_full[comp].detach().to("cpu").numpy().astype(_np.float32)
    NZ_full, NX_full = field.shape
    minv = float(field.min())
    maxv = float(field.max())
    rng = maxv - minv
    if abs(rng) <= 1.0e-12:
        pix = _np.full(field.shape, 128, dtype=_np.uint8)
    else:
        norm = (field.astype(_np.float64) - minv) / rng
        pixf = 1.0 + norm * 254.0
        pix = _np.clip(_np.floor(pixf + 0.5), 1, 255).astype(_np.uint8)
    try:
        with open(filename, "wb") as f:
            f.write(f"P5\n{NX_full} {NZ_full}\n255\n".encode("ascii"))
            f.write(pix.tobytes())
    except OSError as e:
        print(f"[DUMP] fopen failed for {filename!r}: {e}")
        return
    print(f"[DUMP] Wrote {filename} (PGM, {NX_full}x{NZ_full}, "
          f"comp={comp}, min={minv:g}, max={maxv:g})")


# ---------------------------------------------------------------- driver
def run_dns(N=8, Re=100, K0=10.0, STEPS=2, CFL=0.75, backend="auto", start_spectrum="KM3",
            UPDATE=100, method="CNAB2", precision=None):
    """The ``sim`` loop (turbo_simulator.py:2744-2859), same prints and cadence."""
    if method == "CHECK":
        compare_time_steppers(N, Re, K0, STEPS, CFL, backend, start_spectrum, UPDATE, precision)
        return None
    if method not in _METHOD_ID:
        raise ValueError(f"unknown time stepper: {method}")
    print("--- RUN DNS ---")
    print(f" N   = {N}")
    print(f" Re  = {Re}")
    print(f" K0  = {K0}")
    print(f" Steps = {STEPS}")
    print(f" CFL  = {CFL}")
    print(f" Update = {UPDATE}")
    print(f" Start spectrum = {start_spectrum}")
    print(f" Method = {method}")
    print(f" requested = {backend}")
    if backend == "cpu":
        # the CLI's `cpu` (e.g. `sim 256 10000 10 1001 0.75 cpu 100 KM3`): there is
        # no CPU product path here, so the run goes to the GPU and says so
        print(" backend 'cpu' is not built: running on the GPU")
        backend = "gpu"
    seed = pao_seed_from_env(1)
    start = time.perf_counter()
    S = create_dns_state(N=N, Re=Re, K0=K0, CFL=CFL, backend=backend, seed=seed,
                         start_spectrum=start_spectrum, populate_compact_ur=False, precision=precision)
    print(f" effective = gpu (libpalx HIP, {'fp32' if S.precision == 4 else 'fp64'})")
    print(f" DNS INITIALIZATION took {time.perf_counter() - start:.3f} seconds")
    dns_step2a(S)
    CFLM0 = float(compute_cflm(S))
    S.dt = S.cflnum / (CFLM0 * math.pi)
    S.cn = 1.0
    S.cnm1 = 0.0
    S.t = 0.0
    print(f" [NEXTDT INIT] CFLM={CFLM0:11.4f} DT={S.dt:11.7f} CN={S.cn:11.7f}")
    print(f" Initial DT={S.dt:11.7f} CN={S.cn:11.7f}")
    S.sync()
    t0 = time.perf_counter()
    for it in range(1, STEPS + 1):
        S.it = it
        dt_old = S.dt
        if method == "LS_IMEX_RK3":
            dns_step_ls_imex_rk3(S)
        else:
            dns_step2b(S)
            dns_step3(S)
            dns_step2a(S)
        if (it % UPDATE) == 0 or it == 1 or it == STEPS:
            CFLM = next_dt(S, sync_host=True)
            _, _, tau_l = S.flow_eddy_metrics()
            print(f" ITERATION {it:6d} T={S.t:12.10f} DT={S.dt:10.8f} CN={S.cn:10.8f} "
                  f"CFLM={float(CFLM):.6f} TAU_L={tau_l:.6f}")
        S.t += dt_old
    S.sync()
    elap = time.perf_counter() - t0
    fps = (STEPS / elap) if elap > 0 else 0.0
    print(f" Elapsed GPU time for {STEPS} steps (s) = {elap:.8g}")
    print(f" Final T={S.t:.8g}  CN={S.cn:.8g}  DT={S.dt:.8g}")
    print(f" FPS = {fps:.8g}")
    return S


def _relative_errors(reference, candidate) -> tuple[float, float]:
    """rel L2 and rel Linf in complex128 (turbo_simulator.py:2868-2874)."""
    def np_(x):
        return x.detach().cpu().numpy() if hasattr(x, "detach") else _np.asarray(x)
    ref = np_(reference).astype(_np.complex128)
    cand = np_(candidate).astype(_np.complex128)
    d = cand - ref
    l2 = float(_np.linalg.norm(d.ravel()) / max(_np.linalg.norm(ref.ravel()), 1.0e-30))
    linf = float(_np.max(_np.abs(d)) / max(float(_np.max(_np.abs(ref))), 1.0e-30))
    return l2, linf


def compare_time_steppers(N, Re, K0, STEPS, CFL, backend, start_spectrum, UPDATE, precision=None):
    """CHECK: CNAB2 vs LS-IMEX-RK3 from the same seed (turbo_simulator.py:2877-2903).

    Prints the reference's CHECK RESULT lines ("Energy" is U of
    flow_eddy_metrics there too) and returns the printed figures as a dict."""
    print("--- CHECK TIME STEPPERS ---")
    print(" Compare CNAB2 and LS_IMEX_RK3 from the same PAO seed and report final-field errors.")
    a = run_dns(N, Re, K0, STEPS, CFL, backend, start_spectrum, UPDATE, "CNAB2", precision)
    b = run_dns(N, Re, K0, STEPS, CFL, backend, start_spectrum, UPDATE, "LS_IMEX_RK3", precision)
    om_l2, om_linf = _relative_errors(a.om2, b.om2)
    u_l2, u_linf = _relative_errors(a.ur_full[0:2], b.ur_full[0:2])
    e0, _, tau0 = a.flow_eddy_metrics()
    e1, _, tau1 = b.flow_eddy_metrics()
    print("--- CHECK RESULT ---")
    print(f" OM2 relative L2={om_l2:.6e} Linf={om_linf:.6e}")
    print(f" U   relative L2={u_l2:.6e} Linf={u_linf:.6e}")
    print(f" Energy CNAB2={e0:.8e} LS_IMEX_RK3={e1:.8e} diff={abs(e1 - e0):.6e}")
    print(f" Tau_L  CNAB2={tau0:.8e} LS_IMEX_RK3={tau1:.8e} diff={abs(tau1 - tau0):.6e}")
    return {"om_l2": om_l2, "om_linf": om_linf, "u_l2": u_l2, "u_linf": u_linf,
            "e_cnab2": e0, "e_rk3": e1, "tau_cnab2": tau0, "tau_rk3": tau1}


def main(argv=None):
    """``sim N Re K0 STEPS CFL BACKEND UPDATE [SPECTRUM [METHOD]]`` (turbo_simulator.py:2905-2918).

    The precision comes from PALINSTROPHY_PRECISION (fp32 default, fp64)."""
    args = list(sys.argv[1:] if argv is None else argv)
    N = int(args[0]) if len(args) > 0 else 512
    Re = float(args[1]) if len(args) > 1 else 10000
    K0 = float(args[2]) if len(args) > 2 else 10.0
    STEPS = int(args[3]) if len(args) > 3 else 1001
    CFL = float(args[4]) if len(args) > 4 else 0.25
    BACK = cast(Literal["cpu", "gpu", "auto"], args[5].lower()) if len(args) > 5 else "auto"
    UPDATE = int(float(args[6])) if len(args) > 6 else 100
    SPEC = cast(SPECTRUM, args[7].upper()) if len(args) > 7 else "KM3"
    METHOD = cast(TimeStepper, args[8].upper()) if len(args) > 8 else "CNAB2"
    return run_dns(N=N, Re=Re, K0=K0, STEPS=STEPS, CFL=CFL, backend=BACK, start_spectrum=SPEC,
                   UPDATE=UPDATE, method=METHOD)


if __name__ == "__main__":
    main()
