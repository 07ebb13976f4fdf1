"""ctypes binding of libpalx.so (declarations in include/palx.h).

The library is loaded from ``palinstrophy_amd/_lib/libpalx.so`` (built in-tree by
``palinstrophy_amd.build``).  There is no fallback: if the library is missing
or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libpalx.so")

PALX_OK, PALX_EINVAL, PALX_ENOMEM, PALX_EHIP, PALX_ESTATE, PALX_EUNSUP = 0, -1, -2, -3, -4, -5
CNAB2, LS_IMEX_RK3 = 0, 1
STAGE_IN_STEP = 0x1000  # stage_update flag: the LS-IMEX-RK3 stages of one whole step, in order
F_OM, F_FNM1, F_UR, F_WORK, F_UC0, F_SCAL, F_ADJ_OM = 0, 1, 2, 3, 4, 5, 6
S_DT, S_CN, S_CNM1, S_T, S_VISC, S_CFLNUM, S_CFLM, S_USUM, S_INVDX, S_IT, S_NAN = range(11)
NSCAL = 16
SPECTRUM_NOALLOC_MAX_BINS = 32768
VAR_U, VAR_V, VAR_ENERGY, VAR_OMEGA, VAR_STREAM = range(5)

# name -> (restype, argtypes); mirrors include/palx.h
_P = C.c_void_p
_I = C.c_int32
_D = C.c_double
SIGNATURES = {
    "palx_create": (C.c_int, [_I, _I, C.POINTER(_P)]),
    "palx_destroy": (C.c_int, [_P]),
    "palx_set_stream": (C.c_int, [_P, _P]),
    "palx_info": (C.c_int, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "palx_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
    "palx_sync": (C.c_int, [_P]),
    "palx_pao_generate": (C.c_int, [_I, _D, _D, _I, _I, _P, _P]),
    "palx_calcom_f32": (C.c_int, [_I, _P, _P]),
    "palx_load_initial": (C.c_int, [_P, _P, _P]),
    "palx_load_velocity": (C.c_int, [_P]),
    "palx_field_ptr": (C.c_int, [_P, _I, C.POINTER(_P), C.POINTER(C.c_size_t)]),
    "palx_download": (C.c_int, [_P, _I, _P, C.c_size_t]),
    "palx_upload": (C.c_int, [_P, _I, _P, C.c_size_t]),
    "palx_set_scalar": (C.c_int, [_P, _I, _D]),
    "palx_get_scalars": (C.c_int, [_P, _P, _I]),
    "palx_stage_products": (C.c_int, [_P]),
    "palx_stage_update": (C.c_int, [_P, _I, _I]),
    "palx_step": (C.c_int, [_P, _I, _I, _I]),
    "palx_set_iteration": (C.c_int, [_P, C.c_int64]),
    "palx_set_graph": (C.c_int, [_P, _I]),
    "palx_physical": (C.c_int, [_P]),
    "palx_cflm": (C.c_int, [_P, C.POINTER(_D)]),
    "palx_next_dt": (C.c_int, [_P]),
    "palx_diag": (C.c_int, [_P, _P]),
    "palx_spectrum": (C.c_int, [_P, _P, _P, _I]),
    "palx_spectrum_into": (C.c_int, [_P, _P, _P, _I]),
    "palx_spectrum_noalloc": (C.c_int, [_P, _P, _P, _I]),
    "palx_getter": (C.c_int, [_P, _I]),
    "palx_recon_velocity": (C.c_int, [_P]),
    "palx_adj_forward": (C.c_int, [_P, _I, _I]),
    "palx_adj_backward": (C.c_int, [_P, _P]),
    "palx_n_variants": (C.c_int, [_I]),
    "palx_variant_ok": (C.c_int, [_I, _I, _I]),
    "palx_set_variants": (C.c_int, [_P, _I, _I]),
    "palx_get_variants": (C.c_int, [_P, C.POINTER(_I), C.POINTER(_I)]),
    "palx_ens_create": (C.c_int, [_I, _P, _I, C.POINTER(_P)]),
    "palx_ens_destroy": (C.c_int, [_P]),
    "palx_ens_set_stream": (C.c_int, [_P, _P]),
    "palx_ens_member": (C.c_int, [_P, _I, C.POINTER(_P)]),
    "palx_ens_classes": (C.c_int, [_P, C.POINTER(_I), _P, _P, _P, _P]),
    "palx_ens_set_variants": (C.c_int, [_P, _I, _I, _I]),
    "palx_ens_tune": (C.c_int, [_P, _I, _P]),
    "palx_ens_step": (C.c_int, [_P, _I, _I, _I, _I]),
    "palx_ens_set_iteration": (C.c_int, [_P, C.c_int64]),
    "palx_ens_sync": (C.c_int, [_P]),
    "palx_frame_dims": (C.c_int, [_P, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(C.c_size_t)]),
    "palx_frame_reset": (C.c_int, [_P, _D]),
    "palx_frame": (C.c_int, [_P, _I, _I, _I, _P, _P, C.c_size_t]),
}
FRAME_NEXTDT, FRAME_STRETCH, FRAME_ADAPT = 1, 2, 4
(FR_T, FR_DT, FR_CN, FR_VISC, FR_RE, FR_IT, FR_MU, FR_SIGMA, FR_STOP, FR_RATIO, FR_FILT, FR_ADAPTED,
 FR_U, FR_L, FR_TAU_L, FR_EJ, FR_ENSTROPHY, FR_PALINSTROPHY, FR_KMAX2, FR_ENERGY, FR_FMIN, FR_FMAX,
 FR_WIDTH, FR_HEIGHT, FR_CFLM) = range(25)
FRAME_NREC = 32


class FrameCtl(C.Structure):
    """palx_frame_ctl: display stretch and viscosity-controller constants."""
    _fields_ = [("k_std", C.c_double), ("target", C.c_double), ("kp", C.c_double), ("ki", C.c_double),
                ("kd", C.c_double), ("deadband", C.c_double), ("max_frac", C.c_double),
                ("palin_tau", C.c_double), ("re0", C.c_double), ("re_factor", C.c_double)]


MAX_VARIANTS = 8


def tile_width(M: int, extent: int, csz: int) -> int:
    """Tile of the work-plane layouts (tile_width in palx_launch.h): 64 bytes
    of complex data (32 bytes on FP32 grids with M <= 3072), halved until it
    divides `extent`."""
    p = (32 if csz == 8 and M <= 3072 else 64) // csz
    while p > 1 and extent % p:
        p //= 2
    return p


def untile_work(a: np.ndarray) -> np.ndarray:
    """Velocity work planes as downloaded (2, M/P, N/2, P), tiled by z, to
    [c][kx][z] (2, N/2, M)."""
    c, nt, h, p = a.shape
    return np.ascontiguousarray(a.transpose(0, 2, 1, 3).reshape(c, h, nt * p))

ENS_GRAPH, ENS_SERIAL = 1, 2

_lib = None
_lock = threading.Lock()


class PalxError(RuntimeError):
    """A libpalx call returned a nonzero status."""

    def __init__(self, code, msg):
        super().__init__(f"libpalx error {code}: {msg}")
        self.code = code


def load():
    """Load libpalx.so once; raise if it is not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libpalx.so not found at {LIB_PATH}; build it with "
                "`python -m palinstrophy_amd.build` (hipcc, gfx950)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    load().palx_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc: int, what: str = ""):
    if rc != PALX_OK:
        msg = last_error()
        if rc == PALX_EINVAL or rc == PALX_EUNSUP:
            raise ValueError(f"{what}: {msg}")
        raise PalxError(rc, f"{what}: {msg}")


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def pao_generate(N: int, K0: float, Re: float, seed: int, km3: bool):
    """Host PAO/KM3 generator: returns (uc0 complex64 (2,N,N/2), stats float64[16])."""
    uc0 = np.zeros((2, N, N // 2), dtype=np.complex64)
    stats = np.zeros(16, dtype=np.float64)
    check(load().palx_pao_generate(int(N), float(K0), float(Re), int(seed), int(bool(km3)),
                                   _ptr(uc0), _ptr(stats)), "palx_pao_generate")
    return uc0, stats


def calcom_f32(uc0: np.ndarray) -> np.ndarray:
    uc0 = np.ascontiguousarray(uc0, dtype=np.complex64)
    N = uc0.shape[1]
    om = np.zeros((N, N // 2), dtype=np.complex64)
    check(load().palx_calcom_f32(int(N), _ptr(uc0), _ptr(om)), "palx_calcom_f32")
    return om


def n_variants(N: int) -> int:
    """Kernel variants of the step passes compiled for grid size N."""
    n = load().palx_n_variants(int(N))
    if n < 0:
        check(n, "palx_n_variants")
    return n


def variant_ok(N: int, precision: int, v: int) -> bool:
    """Whether step-pass variant v is compiled for grid size N in that
    precision (at N=16384 only the split-exchange variants fit the LDS)."""
    r = load().palx_variant_ok(int(N), int(precision), int(v))
    if r < 0:
        check(r, "palx_variant_ok")
    return r == 1


class Handle:
    """Owns one libpalx simulation handle (device memory + stream).

    ``borrowed`` wraps a handle owned elsewhere (an ensemble member): it is
    never destroyed from here."""

    def __init__(self, N: int, precision: int, borrowed=None):
        self.lib = load()
        if borrowed is not None:
            self.h = borrowed
            self._owned = False
        else:
            h = C.c_void_p()
            check(self.lib.palx_create(int(N), int(precision), C.byref(h)), "palx_create")
            self.h = h
            self._owned = True
        self.N = int(N)
        self.M = 3 * self.N // 2
        self.H = self.N // 2
        self.precision = int(precision)
        self.rdtype = np.float32 if precision == 4 else np.float64
        self.cdtype = np.complex64 if precision == 4 else np.complex128
        self.work_tile = tile_width(self.M, self.M, 2 * precision)

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_owned", True):
                self.lib.palx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def call(self, name, *args):
        check(getattr(self.lib, name)(self.h, *args), name)

    # --- fields
    def shape(self, field):
        N, M, H = self.N, self.M, self.H
        return {F_OM: ((N, H), self.cdtype), F_FNM1: ((N, H), self.cdtype), F_ADJ_OM: ((N, H), self.cdtype),
                F_UR: ((3, M, M), self.rdtype),
                F_WORK: ((2, M // self.work_tile, H, self.work_tile), self.cdtype),
                F_UC0: ((2, N, H), self.cdtype)}[field]

    def download(self, field) -> np.ndarray:
        shp, dt = self.shape(field)
        out = np.empty(shp, dtype=dt)
        self.call("palx_download", field, _ptr(out), out.nbytes)
        return out

    def upload(self, field, arr):
        shp, dt = self.shape(field)
        a = np.ascontiguousarray(arr, dtype=dt).reshape(shp)
        self.call("palx_upload", field, _ptr(a), a.nbytes)

    def upload_ptr(self, field, src_ptr, nbytes):
        """palx_upload from a device (or host) pointer: [z][kx] order for the spectral fields."""
        self.call("palx_upload", field, C.c_void_p(int(src_ptr)), int(nbytes))

    def download_ptr(self, field, dst_ptr, nbytes):
        """palx_download into a device (or host) pointer."""
        self.call("palx_download", field, C.c_void_p(int(dst_ptr)), int(nbytes))

    def device_ptr(self, field):
        p = C.c_void_p()
        n = C.c_size_t()
        self.call("palx_field_ptr", field, C.byref(p), C.byref(n))
        return p.value, n.value

    def set_scalar(self, idx, val):
        self.call("palx_set_scalar", int(idx), float(val))

    def scalars(self) -> np.ndarray:
        out = np.zeros(NSCAL, dtype=np.float64)
        self.call("palx_get_scalars", _ptr(out), NSCAL)
        return out

    def load_initial(self, uc0_c64, om_c64):
        a = np.ascontiguousarray(uc0_c64, dtype=np.complex64)
        b = np.ascontiguousarray(om_c64, dtype=np.complex64)
        self.call("palx_load_initial", _ptr(a), _ptr(b))

    def load_velocity(self, uc0=None):
        """Upload velocity spectra (2, N, N/2) if given, then rebuild the work planes."""
        if uc0 is not None:
            self.upload(F_UC0, uc0)
        self.call("palx_load_velocity")

    def stage_products(self):
        self.call("palx_stage_products")

    def stage_update(self, method, stage=0):
        self.call("palx_stage_update", int(method), int(stage))

    def step(self, nsteps, method, nextdt_every=0):
        self.call("palx_step", int(nsteps), int(method), int(nextdt_every))

    def set_iteration(self, it):
        self.call("palx_set_iteration", C.c_int64(int(it)))

    def set_graph(self, enable=True):
        self.call("palx_set_graph", int(bool(enable)))

    def physical(self):
        self.call("palx_physical")

    def cflm(self) -> float:
        v = C.c_double()
        self.call("palx_cflm", C.byref(v))
        return v.value

    def next_dt(self):
        self.call("palx_next_dt")

    def diag(self) -> np.ndarray:
        out = np.zeros(8, dtype=np.float64)
        self.call("palx_diag", _ptr(out))
        return out

    def spectrum(self, nbins):
        e = np.zeros(nbins, dtype=np.float64)
        c = np.zeros(nbins, dtype=np.int32)
        self.call("palx_spectrum", _ptr(e), _ptr(c), int(nbins))
        return e, c

    def spectrum_noalloc(self, nbins):
        """palx_spectrum without device allocations (static per-GPU staging bins)."""
        e = np.zeros(nbins, dtype=np.float64)
        c = np.zeros(nbins, dtype=np.int32)
        self.call("palx_spectrum_noalloc", _ptr(e), _ptr(c), int(nbins))
        return e, c

    def spectrum_into(self, esum_dev_ptr, count_dev_ptr, nbins):
        """Bin the energy spectrum into device arrays (float64 / int32 pointers)."""
        self.call("palx_spectrum_into", C.c_void_p(int(esum_dev_ptr)), C.c_void_p(int(count_dev_ptr)), int(nbins))

    def recon_velocity(self):
        self.call("palx_recon_velocity")

    def adj_forward(self, nsteps, ckpt_every):
        self.call("palx_adj_forward", int(nsteps), int(ckpt_every))

    def adj_backward(self, lam_T):
        a = np.ascontiguousarray(lam_T, dtype=self.cdtype).reshape(self.N, self.H)
        self.call("palx_adj_backward", _ptr(a))
        return self.download(F_ADJ_OM)

    def getter(self, var):
        self.call("palx_getter", int(var))

    def sync(self):
        self.call("palx_sync")

    def set_variants(self, vrow, vcol):
        self.call("palx_set_variants", int(vrow), int(vcol))

    def frame_dims(self, scale):
        """(width, height, bytes) of a palx_frame output buffer at this scale."""
        w, hg, n = C.c_int32(), C.c_int32(), C.c_size_t()
        self.call("palx_frame_dims", int(scale), C.byref(w), C.byref(hg), C.byref(n))
        return w.value, hg.value, n.value

    def frame_reset(self, re):
        self.call("palx_frame_reset", float(re))

    def frame(self, var, scale, flags, ctl, out_dev_ptr, out_bytes):
        """Enqueue one display frame into the device buffer at out_dev_ptr (out_bytes long)."""
        self.call("palx_frame", int(var), int(scale), int(flags),
                  C.byref(ctl) if ctl is not None else None, C.c_void_p(int(out_dev_ptr)), int(out_bytes))

    def variants(self):
        a, b = C.c_int32(), C.c_int32()
        self.call("palx_get_variants", C.byref(a), C.byref(b))
        return a.value, b.value


class Ensemble:
    """Owns a libpalx ensemble: members of uneven N advanced together, one
    batched launch per pass and size class, one HIP graph launch per step."""

    def __init__(self, Ns, precision: int):
        self.lib = load()
        self.Ns = [int(n) for n in Ns]
        arr = np.ascontiguousarray(self.Ns, dtype=np.int32)
        e = C.c_void_p()
        check(self.lib.palx_ens_create(len(self.Ns), _ptr(arr), int(precision), C.byref(e)), "palx_ens_create")
        self.e = e
        self.precision = int(precision)
        self.members = []
        for m, N in enumerate(self.Ns):
            p = C.c_void_p()
            self.call("palx_ens_member", m, C.byref(p))
            self.members.append(Handle(N, precision, borrowed=p))

    def call(self, name, *args):
        check(getattr(self.lib, name)(self.e, *args), name)

    def close(self):
        if getattr(self, "e", None):
            for h in self.members:
                h.h = None
            self.lib.palx_ens_destroy(self.e)
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def classes(self):
        """[(N, count, vrow, vcol)] in order of first appearance."""
        n = len(self.Ns)
        N, cnt, vr, vc = (np.zeros(n, dtype=np.int32) for _ in range(4))
        nc = C.c_int32()
        self.call("palx_ens_classes", C.byref(nc), _ptr(N), _ptr(cnt), _ptr(vr), _ptr(vc))
        return [(int(N[i]), int(cnt[i]), int(vr[i]), int(vc[i])) for i in range(nc.value)]

    def set_variants(self, N, vrow, vcol):
        self.call("palx_ens_set_variants", int(N), int(vrow), int(vcol))

    def tune(self, reps=10):
        """Time every variant of every class; keep the fastest.  Overwrites the
        member fields: call before loading initial fields.  Returns
        ms[class][pass][variant] (-1 where a variant does not exist)."""
        nc = len(self.classes())
        ms = np.zeros((nc, 2, MAX_VARIANTS), dtype=np.float32)
        self.call("palx_ens_tune", int(reps), _ptr(ms))
        return ms

    def step(self, nsteps, method, nextdt_every=0, graph=True, serial=False):
        flags = (ENS_GRAPH if graph else 0) | (ENS_SERIAL if serial else 0)
        self.call("palx_ens_step", int(nsteps), int(method), int(nextdt_every), flags)

    def set_iteration(self, it):
        self.call("palx_ens_set_iteration", C.c_int64(int(it)))

    def sync(self):
        self.call("palx_ens_sync")
