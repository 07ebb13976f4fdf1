"""Restart files in the reference's layout (``restart.nc``).

The reference GUI writes its restart with netCDF4 in NETCDF4 (HDF5) format
(turbo_main.py:1283-1338) and reads it back in _load_restart_netcdf
(turbo_main.py:1373-1503).  Neither netCDF4 nor an HDF5 library is part of
this stack, so the file is written here as NetCDF-3 (64-bit offset) through
scipy.io with the same global attributes, dimension names and variables
(``<name>_real`` / ``<name>_imag`` float32 pairs for uc, om2, fnm1).  The
reference's reader opens such a file unchanged (netCDF4.Dataset reads the
classic formats); this module reads the files it writes, and refuses an
HDF5-based file with a clear message.

``uc`` (N, 3N/4+1, 3) holds the velocity spectra reconstructed from the
current omega (turbo_simulator.py:1555-1585) in components 0/1 and zeros in
component 2, so that a reader that rebuilds the physical field from ``uc``
gets the current flow, not the initial one.  Loading rebuilds the velocity
from omega on the device (palx_recon_velocity).
"""
from __future__ import annotations

import os

import numpy as np

RESTART_FILE = "restart.nc"
FORMAT_VERSION = 1
_HDF5_MAGIC = b"\x89HDF\r\n\x1a\n"

_ATTRS = ("format_version", "Nbase", "Re", "K0", "start_spectrum", "visc", "cflnum", "seed_init",
          "t", "dt", "cn", "cnm1", "it")


def _dims(name: str, ndim: int) -> tuple[str, ...]:
    """turbo_main.py:1283-1291 _restart_dims."""
    if name == "uc":
        return ("uc_z", "uc_k", "uc_component")
    if name == "om2":
        return ("om2_z", "om2_x_half")
    if name == "fnm1":
        return ("fnm1_z", "fnm1_x_half")
    return tuple(f"{name}_dim_{i}" for i in range(ndim))


def velocity_from_omega(om: np.ndarray) -> np.ndarray:
    """Compact uc (N, 3N/4+1, 3) complex64 from omega (N, N/2):
    u = -i kz om/(K2+1e-30), w = i kx om/(K2+1e-30) for kx >= 1; for kx = 0
    u = -i om/kz (0 at kz = 0), w = 0 (turbo_simulator.py:1555-1585)."""
    N, H = om.shape
    NK = 3 * N // 4 + 1
    om = om.astype(np.complex128)
    kx = np.arange(H, dtype=np.float64)[None, :]
    kz = np.array([z if z < N // 2 else z - N for z in range(N)], dtype=np.float64)[:, None]
    K2 = kx * kx + kz * kz + 1.0e-30
    u = -1j * kz * om / K2
    w = 1j * kx * om / K2
    kz0 = kz[:, 0]
    with np.errstate(divide="ignore", invalid="ignore"):
        u[:, 0] = np.where(kz0 != 0.0, -1j * om[:, 0] / np.where(kz0 != 0.0, kz0, 1.0), 0.0)
    w[:, 0] = 0.0
    uc = np.zeros((N, NK, 3), dtype=np.complex64)
    uc[:, :H, 0] = u
    uc[:, :H, 1] = w
    return uc


def write_restart_file(path: str, attrs: dict, arrays: dict, f8=()) -> None:
    """Write attributes and complex arrays in the reference's variable layout.
    Arrays named in `f8` also get float64 pairs ``<name>_real_f8`` /
    ``<name>_imag_f8`` (an FP64 state keeps its precision; the reference's
    reader only looks at the float32 pairs)."""
    from scipy.io import netcdf_file

    missing = [a for a in _ATTRS if a not in attrs]
    if missing:
        raise ValueError(f"restart attributes missing: {missing}")
    with netcdf_file(path, "w", version=2) as ds:
        for a in _ATTRS:
            v = attrs[a]
            if isinstance(v, str):
                setattr(ds, a, v)
            elif isinstance(v, (int, np.integer)):
                setattr(ds, a, np.int32(v))
            else:  # doubles stay doubles (scipy would store a bare float as float32)
                setattr(ds, a, np.array([float(v)], dtype=np.float64))
        for name, arr in arrays.items():
            host = np.asarray(arr)
            dims = _dims(name, host.ndim)
            for d, n in zip(dims, host.shape):
                ds.createDimension(d, int(n))
            re = ds.createVariable(f"{name}_real", "f4", dims)
            im = ds.createVariable(f"{name}_imag", "f4", dims)
            re[:] = np.asarray(host.real, dtype=np.float32)
            im[:] = np.asarray(host.imag, dtype=np.float32)
            if name in f8:
                re8 = ds.createVariable(f"{name}_real_f8", "f8", dims)
                im8 = ds.createVariable(f"{name}_imag_f8", "f8", dims)
                re8[:] = np.asarray(host.real, dtype=np.float64)
                im8[:] = np.asarray(host.imag, dtype=np.float64)


def read_restart_file(path: str) -> tuple[dict, dict]:
    """(attrs, {uc, om2, fnm1}) of a restart file written by write_restart_file;
    complex128 where float64 pairs are present, complex64 otherwise."""
    from scipy.io import netcdf_file

    with open(path, "rb") as f:
        head = f.read(8)
    if head == _HDF5_MAGIC:
        raise RuntimeError(f"{path} is a NETCDF4/HDF5 restart (the reference GUI's format); reading it "
                           "needs the netCDF4 package, which this installation does not have")
    attrs, arrays = {}, {}
    with netcdf_file(path, "r", mmap=False) as ds:
        for a in _ATTRS:
            v = getattr(ds, a)
            if isinstance(v, bytes):
                attrs[a] = v.decode()
            elif np.issubdtype(np.asarray(v).dtype, np.integer):
                attrs[a] = int(np.asarray(v).reshape(-1)[0])
            else:
                attrs[a] = float(np.asarray(v).reshape(-1)[0])
        for name in ("uc", "om2", "fnm1"):
            if f"{name}_real_f8" in ds.variables:
                re = np.array(ds.variables[f"{name}_real_f8"][:], dtype=np.float64)
                im = np.array(ds.variables[f"{name}_imag_f8"][:], dtype=np.float64)
                arrays[name] = re + 1j * im
            else:
                re = np.array(ds.variables[f"{name}_real"][:], dtype=np.float32)
                im = np.array(ds.variables[f"{name}_imag"][:], dtype=np.float32)
                arrays[name] = (re + 1j * im).astype(np.complex64)
    return attrs, arrays


def dump_restart_netcdf(S, folder_path: str) -> str:
    """Write ``folder_path/restart.nc`` for DnsState S (turbo_main.py:1312-1338)."""
    from . import _palx as P
    from .turbo_simulator import sync_time_scalars_from_device

    sync_time_scalars_from_device(S)
    om = S.handle.download(P.F_OM)
    fn = S.handle.download(P.F_FNM1)
    attrs = {"format_version": FORMAT_VERSION, "Nbase": int(S.Nbase), "Re": float(S.Re), "K0": float(S.K0),
             "start_spectrum": str(S.start_spectrum), "visc": float(S.visc), "cflnum": float(S.cflnum),
             "seed_init": int(S.seed_init), "t": float(S.t), "dt": float(S.dt), "cn": float(S.cn),
             "cnm1": float(S.cnm1), "it": int(S.it)}
    os.makedirs(folder_path, exist_ok=True)
    path = os.path.join(folder_path, RESTART_FILE)
    f8 = ("om2", "fnm1") if om.dtype == np.complex128 else ()
    write_restart_file(path, attrs, {"uc": velocity_from_omega(om), "om2": om, "fnm1": fn}, f8=f8)
    print(f"[SAVE] NetCDF restart written to {path}")
    return path


def load_restart_netcdf(folder_path: str, backend: str = "gpu", precision: str | None = None):
    """Recreate a DnsState from ``folder_path/restart.nc`` (turbo_main.py:1373-1503):
    a state of the stored N without the seeded field, omega and fnm1 loaded,
    the scalars restored, the velocity rebuilt from omega."""
    from . import _palx as P
    from .turbo_simulator import create_dns_state, dns_load_spectral

    path = os.path.join(folder_path, RESTART_FILE)
    if not os.path.isfile(path):
        raise FileNotFoundError(f"no {RESTART_FILE} in {folder_path}")
    attrs, arrays = read_restart_file(path)
    if int(attrs["format_version"]) != FORMAT_VERSION:
        raise ValueError(f"unsupported restart format_version {attrs['format_version']}")
    N = int(attrs["Nbase"])
    S = create_dns_state(N=N, Re=float(attrs["Re"]), K0=float(attrs["K0"]), CFL=float(attrs["cflnum"]),
                         backend=backend, seed=int(attrs["seed_init"]), skip_pao=True,
                         start_spectrum=str(attrs["start_spectrum"]), precision=precision)
    S.visc = float(attrs["visc"])
    S.cflnum = float(attrs["cflnum"])
    S.t = float(attrs["t"])
    S.dt = float(attrs["dt"])
    S.cn = float(attrs["cn"])
    S.cnm1 = float(attrs["cnm1"])
    S.it = int(attrs["it"])
    S.handle.set_scalar(P.S_T, S.t)
    S.handle.set_iteration(S.it)
    dns_load_spectral(S, arrays["om2"], arrays["fnm1"])
    print(f"[LOAD] Restored case from {folder_path}: N={N}, Re={S.Re:.4e}, K0={S.K0}, "
          f"spectrum={S.start_spectrum}, t={S.t:.6e}, it={S.it}")
    return S
