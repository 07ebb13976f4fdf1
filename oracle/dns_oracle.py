"""NumPy/scipy.fft restatement of the reference DNS CPU path (test oracle only).

Every function follows the CPU branch of ``palinstrophy/turbo_simulator.py``
(cited as ``ts.py:LINE``) op for op, in the working precision ``dtype``
(float32 = the reference's own precision, float64 = the FP64 restatement).

Conventions shared with the HIP implementation (and documented there):
  * The seeded initial field is always built in FP32/complex64, exactly as the
    reference does.  Transcendentals of float32 arguments (cos, sin, exp) are
    taken as the correctly rounded float32 of the double-precision function
    (``_f32_of``); the reference used Numba/libm for these, which is
    within 1 ulp of the same values.
  * The FP64 restatement starts from the complex64 seeded spectra and the
    complex64 omega_0 (computed in FP32 like ``ts.py:1351-1361``), promoted
    to complex128.  ``visc`` is the reference's float32(1/Re).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np
import scipy.fft as spfft

IMM, IT, ID = 420029, 2017, 5011          # ts.py:242-244
PI_F32 = np.float32(3.14159265358979)      # ts.py:237
AMP2_FLOOR = np.float32(1.0e-40)           # ts.py:263
RK3_ALPHA = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)    # ts.py:1902
RK3_BETA = (0.0, -17.0 / 60.0, -5.0 / 12.0)        # ts.py:1903


def _f32_of(fn, x32):
    """Correctly rounded float32 of a double-precision function of float32 input."""
    return fn(np.asarray(x32, dtype=np.float64)).astype(np.float32)


def lcg_stream(seed: int, count: int):
    """RANVEC warm-up + fill, then ``count`` shuffled draws (ts.py:242-293).

    Returns (v_seeds, seed_out): the LCG state whose value r=seed/ID was picked
    from the 97-entry shuffle table for each draw, in loop order.
    """
    rtab = (np.arange(ID, dtype=np.float32) / np.float32(ID)).astype(np.float32)
    idxtab = [min(max(int(float(r) * 97.0), 0), 96) for r in rtab]
    s = int(seed)
    for _ in range(97):                               # ts.py:254-255
        s = (s * IMM + IT) % ID
    ranvec = []
    for _ in range(97):                               # ts.py:258-260
        s = (s * IMM + IT) % ID
        ranvec.append(s)
    out = np.empty(count, dtype=np.int64)
    for n in range(count):                            # ts.py:280-290
        s = (s * IMM + IT) % ID
        i = idxtab[s]
        out[n] = ranvec[i]
        ranvec[i] = s
    return out, s, rtab


def hash01(x, z, seed, salt) -> float:
    """Sin-hash used for KM3 jitter and phases (ts.py:195-204)."""
    a = ((float(x) + 1.0) * 12.9898 + (float(z) + 1.0) * 78.233
         + (float(seed) + 1.0) * 0.010001 + (float(salt) + 1.0) * 37.719)
    s = math.sin(a) * 43758.5453123
    return s - math.floor(s)


def wavenumbers(N: int):
    """ALFA(N/2), GAMMA(N) as float32 (ts.py:1079-1094)."""
    ND2 = N // 2
    alfa = np.arange(ND2, dtype=np.float32)
    gamma = np.zeros(N, dtype=np.float32)
    for z in range(1, ND2 + 1):
        gamma[z] = np.float32(z)
        gamma[N - z] = -gamma[z]
    return alfa, gamma


@dataclass
class PaoResult:
    UR: np.ndarray          # complex64 (N/2, N, 2): [x, z, comp]
    seed_out: int
    visc: np.float32
    Q2: float
    W2: float
    E110: float
    A: tuple


def pao_build(N: int, K0: float, Re: float, seed: int, km3: bool) -> PaoResult:
    """Seeded PAO / KM3 spectral velocity field (ts.py:213-485)."""
    ND2 = N // 2
    NE = N
    NED2 = NE // 2
    K0 = np.float32(K0)
    alfa, gamma = wavenumbers(N)

    vseeds, seed_out, rtab = lcg_stream(seed, NE * NED2)
    v = rtab[vseeds].reshape(NE, NED2)                # [z, x] loop order
    th = (np.float32(2.0) * PI_F32) * v               # ts.py:292
    c = _f32_of(np.cos, th)
    s = _f32_of(np.sin, th)

    ax = alfa[None, :].astype(np.float32)             # (1, ND2)
    gz = gamma[:, None].astype(np.float32)            # (NE, 1)
    K2 = (ax * ax + gz * gz).astype(np.float32)       # ts.py:296
    K = np.where(K2 > 0, np.sqrt(K2), np.float32(0)).astype(np.float32)
    NORM = PI_F32 * K0 * K0                           # ts.py:262
    q = (K / K0).astype(np.float32)
    e = _f32_of(np.exp, (-q) * q)

    with np.errstate(divide="ignore", invalid="ignore"):
        absu2 = (e / NORM).astype(np.float32)                           # ts.py:303
        denom = (np.float32(1.0) + (gz * gz) / (ax * ax)).astype(np.float32)  # ts.py:309
        absw2 = (e / (denom * NORM)).astype(np.float32)                 # ts.py:310
        ratio = (-(gz / ax)).astype(np.float32)                         # ts.py:316
    if km3:
        absu2 = np.where(absu2 < AMP2_FLOOR, AMP2_FLOOR, absu2)
        absw2 = np.where(absw2 < AMP2_FLOOR, AMP2_FLOOR, absw2)
    amp = np.sqrt(absu2).astype(np.float32)
    ampw = np.sqrt(absw2).astype(np.float32)

    U0 = np.zeros((NE, ND2), dtype=np.complex64)
    U1 = np.zeros((NE, ND2), dtype=np.complex64)
    # ax == 0 column: purely u1 mode (ts.py:299-307)
    U0[:, 0] = (amp[:, 0] * c[:, 0]) + 1j * (amp[:, 0] * s[:, 0])
    # ax != 0: w = ampw*ARG, u = -(gz/ax)*w (ts.py:309-319)
    wr = (ampw[:, 1:] * c[:, 1:]).astype(np.float32)
    wi = (ampw[:, 1:] * s[:, 1:]).astype(np.float32)
    r = np.broadcast_to(ratio[:, 1:], wr.shape)
    U1[:, 1:] = wr + 1j * wi
    U0[:, 1:] = (r * wr).astype(np.float32) + 1j * (r * wi).astype(np.float32)

    UR = np.zeros((ND2, NE, 2), dtype=np.complex64)
    UR[:, :, 0] = U0.T
    UR[:, :, 1] = U1.T
    UR[0, 0, :] = 0                                   # ts.py:322-323
    for z in range(1, NED2):                          # ts.py:328-330
        UR[0, NE - z, 0] = np.conj(UR[0, z, 0])
        UR[0, NE - z, 1] = np.conj(UR[0, z, 1])
    UR[:, NED2, :] = 0                                # ts.py:333-335

    if km3:
        _km3_rescale(UR, alfa, gamma, N, float(K0), int(seed))

    # averages (ts.py:431-467): sequential double accumulation, x outer / z inner
    u1 = UR[:, :, 0]
    u3 = UR[:, :, 1]
    u1u1 = u1.real.astype(np.float64) ** 2 + u1.imag.astype(np.float64) ** 2
    u3u3 = u3.real.astype(np.float64) ** 2 + u3.imag.astype(np.float64) ** 2
    ax2 = (alfa.astype(np.float64) ** 2)[:, None]
    gz2 = (gamma.astype(np.float64) ** 2)[None, :]
    m = np.full((ND2, 1), 2.0)
    m[0, 0] = 1.0

    def seqsum(a):
        return float(np.cumsum(np.broadcast_to(a, (ND2, NE)).ravel())[-1])

    A1 = seqsum(m * u1u1)
    A2 = seqsum(m * u3u3)
    A3 = seqsum(m * u1u1 * ax2)
    A4 = seqsum(m * u1u1 * gz2)
    A5 = seqsum(m * u3u3 * ax2)
    A6 = seqsum(m * u3u3 * gz2)
    K2f = ax2 + gz2
    A7 = seqsum(m * (u1u1 + u3u3) * K2f * K2f)
    E110 = float(np.cumsum(u1u1[0, :])[-1])
    Q2 = A1 + A2
    W2 = A3 + A4 + A5 + A6
    visc = np.float32(1.0 / float(np.float32(Re)))    # ts.py:470, 485
    return PaoResult(UR, seed_out, visc, Q2, W2, E110, (A1, A2, A3, A4, A5, A6, A7))


def _km3_rescale(UR, alfa, gamma, N, K0f, seed_init):
    """Shell-rescale to the k^-3 start spectrum (ts.py:337-426)."""
    ND2 = N // 2
    NE = N
    NED2 = NE // 2
    k_nyq = float(ND2)
    kmax = ND2
    xs = np.repeat(np.arange(ND2), NE)                # x outer
    zs = np.tile(np.arange(NE), ND2)                  # z inner
    ax = alfa.astype(np.float64)[xs]
    gz = gamma.astype(np.float64)[zs]
    ax2 = ax * ax
    gz2 = gz * gz
    k = np.sqrt(ax2 + gz2)
    shell = (k + 0.5).astype(np.int64)
    inside = (k > 0.0) & (k <= k_nyq) & (shell <= kmax)
    mult = np.where(xs == 0, 1.0, 2.0)
    U1 = UR[xs, zs, 0]
    U3 = UR[xs, zs, 1]
    e = ((U1.real.astype(np.float64) ** 2 + U1.imag.astype(np.float64) ** 2)
         + (U3.real.astype(np.float64) ** 2 + U3.imag.astype(np.float64) ** 2))
    neg_col = (xs == 0) & (zs > NED2)
    zh = np.where(neg_col, NE - zs, zs)
    idx = np.nonzero(inside)[0]
    jit = np.array([0.75 + 0.50 * hash01(xs[i], zh[i], seed_init, 17) for i in idx])
    shell_energy = np.zeros(kmax + 1)
    shell_weight = np.zeros(kmax + 1)
    np.add.at(shell_energy, shell[idx], mult[idx] * e[idx])
    np.add.at(shell_weight, shell[idx], mult[idx] * jit)

    k0f = max(K0f, 1.0)
    target = np.zeros(kmax + 1)
    for sh in range(1, kmax + 1):
        kk = float(sh)
        if kk < k0f:
            low = kk / k0f
            target[sh] = (low * low * low * low) / (k0f * k0f * k0f)
        else:
            target[sh] = 1.0 / (kk * kk * kk)
    orig = 0.0
    tgt = 0.0
    for sh in range(1, kmax + 1):
        orig += shell_energy[sh]
        tgt += target[sh]
    if not (orig > 0.0 and tgt > 0.0):
        return
    norm = orig / tgt
    keep = inside & (shell_weight[np.minimum(shell, kmax)] > 0.0)
    newU1 = np.zeros(ND2 * NE, dtype=np.complex64)
    newU3 = np.zeros(ND2 * NE, dtype=np.complex64)
    for i in np.nonzero(keep)[0]:
        x = int(xs[i])
        sign = -1.0 if neg_col[i] else 1.0
        z_h = int(zh[i])
        jitter = 0.75 + 0.50 * hash01(x, z_h, seed_init, 17)
        me = (norm * target[shell[i]]) * jitter / shell_weight[shell[i]]
        th = sign * 2.0 * math.pi * hash01(x, z_h, seed_init, 53)
        pr = np.float32(math.cos(th))
        pi_ = np.float32(math.sin(th))
        if ax[i] == 0.0:
            a = np.float32(math.sqrt(me))
            newU1[i] = complex(a * pr, a * pi_)
        else:
            denom = 1.0 + gz2[i] / ax2[i]
            a = np.float32(math.sqrt(me / denom))
            wr, wi = a * pr, a * pi_
            r = np.float32(-gz[i] / ax[i])
            newU1[i] = complex(r * wr, r * wi)
            newU3[i] = complex(wr, wi)
    UR[:, :, 0] = newU1.reshape(ND2, NE)
    UR[:, :, 1] = newU3.reshape(ND2, NE)


def re_from_n_k0(N, K0):
    """Re rule log10 Re = a log10 N + b log10 K0 + c (turbo_main.py:2323-2328)."""
    return 10.0 ** (1.066240 * np.log10(N) - 0.083911 * np.log10(K0) + 1.286396)


@dataclass
class OracleState:
    N: int
    dtype: type
    Re: float
    K0: float
    cflnum: float
    spectrum: str
    seed: int
    workers: int
    visc: float = 0.0
    t: float = 0.0
    dt: float = 0.0
    cn: float = 1.0
    cnm1: float = 0.0
    it: int = 0
    pao: PaoResult | None = None
    arrays: dict = field(default_factory=dict)

    def __getattr__(self, name):
        arrays = self.__dict__.get("arrays")
        if arrays is not None and name in arrays:
            return arrays[name]
        raise AttributeError(name)

    @property
    def M(self):
        return 3 * self.N // 2

    @property
    def H(self):
        return self.N // 2

    @property
    def cdtype(self):
        return np.complex64 if self.dtype == np.float32 else np.complex128


def create_state(N, Re, K0, CFL, seed=1, spectrum="KM3", dtype=np.float64,
                 workers=None, uc0=None):
    """create_dns_state + dns_pao_host_init + calcom (ts.py:878-1054, 1059-1239).

    ``uc0`` (optional, complex64 (2, N, N/2)) replaces the seeded spectra, for
    feeding an externally generated initial field.
    """
    dt_ = np.dtype(dtype).type
    S = OracleState(N=N, dtype=dt_, Re=float(Re), K0=float(K0), cflnum=float(CFL),
                    spectrum=spectrum, seed=int(seed),
                    workers=workers or len(os.sched_getaffinity(0)))
    M, H, Mh = S.M, S.H, S.M // 2 + 1
    cd = S.cdtype
    alfa, gamma = wavenumbers(N)
    if uc0 is None:
        S.pao = pao_build(N, K0, Re, seed, spectrum == "KM3")
        UR = S.pao.UR
        uc0 = np.stack([UR[:, :, 0].T, UR[:, :, 1].T]).astype(np.complex64)
        S.visc = float(S.pao.visc)
    else:
        S.visc = float(np.float32(1.0 / float(np.float32(Re))))
    uc_full = np.zeros((3, M, Mh), dtype=cd)
    uc_full[0:2, :N, :H] = uc0                         # ts.py:1205-1225
    # calcom in FP32 (ts.py:1314-1361), then promoted
    gz32 = gamma[:, None]
    ax32 = alfa[None, :]
    u1 = uc0[0]
    u2 = uc0[1]
    dr = (gz32 * u1.real).astype(np.float32) - (ax32 * u2.real).astype(np.float32)
    di = (gz32 * u1.imag).astype(np.float32) - (ax32 * u2.imag).astype(np.float32)
    om2 = ((-di) + 1j * dr).astype(np.complex64).astype(cd)

    ax = alfa[None, :].astype(dt_)
    gz = gamma[:, None].astype(dt_)
    K2 = (ax * ax + gz * gz).astype(dt_)
    S.arrays = dict(
        alfa=alfa.astype(dt_), gamma=gamma.astype(dt_),
        uc0=uc0,
        ur_full=np.zeros((3, M, M), dtype=dt_),
        uc_full=uc_full,
        om2=om2,
        fnm1=np.zeros((N, H), dtype=cd),
        K2=K2,
        GA=(gz * ax).astype(dt_),
        G2mA2=(gz * gz - ax * ax).astype(dt_),
        invK2_sub=(dt_(1.0) / (K2[:, 1:] + dt_(1.0e-30))).astype(dt_),
        inv_gamma0=np.where((np.arange(N) >= 1) & (np.abs(gamma) > 0),
                            dt_(1.0) / np.where(np.abs(gamma) > 0, gamma.astype(dt_), dt_(1.0)),
                            dt_(0.0)).astype(dt_),
    )
    NX32 = dt_(1.5) * dt_(N)
    S.arrays["divxz"] = dt_(dt_(1.0) / (NX32 * NX32))  # ts.py:1050-1052
    S.inv_dx = float(N) / (2.0 * math.pi)              # ts.py:940
    return S


def step2a(S):
    """Dealias, row move, unnormalised irfft2 (ts.py:2020-2037, 1246-1262)."""
    N, M = S.N, S.M
    UC = S.uc_full
    Mh = M // 2 + 1
    hi_end = min(3 * N // 4, Mh - 1)
    UC[0:2, :, N // 2:hi_end + 1] = 0
    h = N // 2
    UC[0:2, N:N + h, :h] = UC[0:2, h:N, :h]
    UC[0:2, h:N, :h] = 0
    ur01 = spfft.irfft2(UC[0:2], s=(M, M), axes=(1, 2), workers=S.workers, norm="forward")
    S.ur_full[0:2] = ur01


def step2b(S):
    """Products in place, rfft2, zero row z=N (ts.py:1467-1473, 1512-1513)."""
    UR = S.ur_full
    u, w = UR[0], UR[1]
    np.multiply(u, w, out=UR[2])
    np.multiply(u, u, out=UR[0])
    np.multiply(w, w, out=UR[1])
    S.uc_full[...] = spfft.rfft2(UR, axes=(1, 2), workers=S.workers)
    N = S.N
    S.uc_full[0:3, N, 0:N // 2] = 0


def nonlinear(S):
    """fn = ((F[u^2]-F[w^2])*GA + F[uw]*(G2-A2)) * divxz (ts.py:1520-1552)."""
    N, H = S.N, S.H
    h = N // 2
    uc = S.uc_full[:, :, :H]
    th = np.concatenate([uc[:, :h], uc[:, N:N + h]], axis=1)   # (3, N, H)
    out = (th[0] - th[1]) * S.GA
    out = out + th[2] * S.G2mA2
    return (out * S.divxz).astype(S.cdtype)


def reconstruct(S):
    """u_hat, w_hat from om2 into uc_full[0:2, :N, :H] (ts.py:1555-1585)."""
    N, H = S.N, S.H
    om2 = S.om2
    cd = S.cdtype
    out1 = np.zeros((N, H), dtype=cd)
    out2 = np.zeros((N, H), dtype=cd)
    if H > 1:
        o1 = om2[:, 1:] * S.invK2_sub
        o1 = o1 * S.gamma[:, None]
        out1[:, 1:] = o1 * cd(-1j)
        o2 = om2[:, 1:] * S.invK2_sub
        o2 = o2 * S.alfa[1:][None, :]
        out2[:, 1:] = o2 * cd(1j)
    out1[:, 0] = cd(-1j) * om2[:, 0] * S.inv_gamma0
    out2[:, 0] = 0
    S.uc_full[0, :N, :H] = out1
    S.uc_full[1, :N, :H] = out2


def step3(S):
    """CNAB2 update (ts.py:1731-1771)."""
    f = S.dtype
    visc = f(S.visc)
    dt = f(S.dt)
    cnm1 = f(S.cnm1)
    fn = nonlinear(S)
    VT = f(0.5) * visc * dt
    ARG = (S.K2 * VT).astype(f)
    DEN = (ARG + f(1.0)).astype(f)
    c2 = f(0.5) * dt * (f(2.0) + cnm1)
    c3 = -f(0.5) * dt * cnm1
    NUM = S.om2 - S.om2 * ARG
    NUM = NUM + fn * c2
    NUM = NUM + S.fnm1 * c3
    S.arrays["om2"] = (NUM / DEN).astype(S.cdtype)
    S.arrays["fnm1"] = fn
    reconstruct(S)
    S.cnm1 = float(S.cn)


def rk3(S):
    """LS-IMEX-RK3 step, three stages of 2B -> update -> 2A (ts.py:1912-1943)."""
    f = S.dtype
    dt = f(S.dt)
    visc = f(S.visc)
    for stage in range(3):
        a = f(RK3_ALPHA[stage])
        b = f(RK3_BETA[stage])
        step2b(S)
        fn = nonlinear(S)
        tmp = (S.K2 * (b * dt * visc)).astype(f)
        tmp = (f(1.0) - tmp).astype(f)
        num = S.om2 * tmp
        num = num + fn * (a * dt)
        num = num + S.fnm1 * (b * dt)
        den = (S.K2 * (a * dt * visc)).astype(f)
        den = (den + f(1.0)).astype(f)
        S.arrays["om2"] = (num / den).astype(S.cdtype)
        S.arrays["fnm1"] = fn
        reconstruct(S)
        step2a(S)


def cnab2(S):
    step2b(S)
    step3(S)
    step2a(S)


def compute_cflm(S):
    """max(|u|+|w|) over the padded grid times N/2pi (ts.py:2106-2125)."""
    m = np.max(np.abs(S.ur_full[0]) + np.abs(S.ur_full[1]))
    if S.dtype == np.float32:
        return float(np.float32(m) * np.float32(S.inv_dx))
    return float(m) * S.inv_dx


def next_dt(S):
    """CFL-adaptive dt (ts.py:2127-2167)."""
    CFLM = compute_cflm(S)
    if CFLM <= 0.0 or S.dt <= 0.0:
        return CFLM
    CFL = CFLM * S.dt * math.pi
    S.cn = 0.8 + 0.2 * (S.cflnum / CFL)
    S.dt = S.dt * S.cn
    return CFLM


def init_dt(S):
    """Initial STEP2A + CFL dt (ts.py:2812-2823)."""
    step2a(S)
    CFLM0 = compute_cflm(S)
    S.dt = S.cflnum / (CFLM0 * math.pi)
    S.cn = 1.0
    S.cnm1 = 0.0
    S.t = 0.0
    return CFLM0


def rfft_weight(H, dtype):
    w = np.full(H, 2.0, dtype=dtype)
    w[0] = 1.0
    if H > 1:
        w[-1] = 1.0
    return w


def flow_eddy_metrics(S):
    """U, L, tau_L (ts.py:764-810); elementwise in working precision, f64 sums."""
    f = S.dtype
    u, v = S.ur_full[0], S.ur_full[1]
    U = math.sqrt(max(0.0, float(np.mean(u * u + v * v, dtype=np.float64))))
    k2 = S.K2
    w = rfft_weight(k2.shape[1], f)
    nz = k2 > 0
    inv_k = np.where(nz, f(1.0) / np.sqrt(np.where(nz, k2, f(1.0))), f(0.0)).astype(f)
    p = (S.om2.real * S.om2.real + S.om2.imag * S.om2.imag).astype(f)
    power = np.where(nz, p / np.where(nz, k2, f(1.0)), f(0.0)).astype(f)
    wp = (power * w).astype(f)
    esum = float(np.sum(wp, dtype=np.float64) - wp[0, 0])
    if esum <= 0.0:
        return U, float("nan"), float("nan")
    lsum = float(np.sum((wp * inv_k).astype(f), dtype=np.float64))
    L = 2.0 * math.pi * lsum / esum
    return U, L, (L / U if U > 0.0 else float("nan"))


def pal_over_ens_kmax2(S):
    """Palinstrophy / (enstrophy * max K2) (turbo_main.py:2015-2073)."""
    f = S.dtype
    w = rfft_weight(S.H, f)
    P = (S.om2.real * S.om2.real + S.om2.imag * S.om2.imag).astype(f)
    total = float(np.sum((w * P).astype(f), dtype=np.float64)) - float(P[0, 0])
    pal = float(np.sum((w * S.K2 * P).astype(f), dtype=np.float64))
    kmax2 = float(S.K2.max())
    if total <= 0.0 or kmax2 <= 0.0:
        return 0.0, total, pal, kmax2
    return pal / (total * kmax2), total, pal, kmax2


def energy_joules(S):
    """0.5 (2pi)^2 sum(u^2+v^2) / M^2 (turbo_main.py:1947-1959)."""
    u, v = S.ur_full[0], S.ur_full[1]
    s = float(np.sum(u * u, dtype=np.float64) + np.sum(v * v, dtype=np.float64))
    return 0.5 * (2.0 * math.pi) ** 2 * s / u.size


def spectrum_bins(S):
    """Radial energy-spectrum bins from om2 (ts.py:2317-2383)."""
    n = S.N
    nbins = max(32, 2 * n)
    r_max = math.sqrt(2.0)
    k_nyq = 0.5 * float(n)
    nx_half = n // 2
    fft_scale2 = float(S.M * S.M) ** 2
    esum = np.zeros(nbins)
    count = np.zeros(nbins, dtype=np.int64)
    kx = np.arange(nx_half, dtype=np.float64)
    kx2 = kx * kx
    weight = np.ones(nx_half)
    if nx_half > 1:
        weight[1:] = 2.0
    om2 = S.om2
    for z in range(n):
        kz = z if z < n // 2 else z - n
        k2 = kx2 + float(kz * kz)
        mask = k2 > 0.0
        radii = np.sqrt(k2[mask]) / k_nyq
        bins = np.floor((radii / r_max) * float(nbins)).astype(np.int64)
        np.clip(bins, 0, nbins - 1, out=bins)
        om = om2[z, mask]
        p = (om.real.astype(np.float64) ** 2 + om.imag.astype(np.float64) ** 2) / k2[mask]
        rw = weight[mask]
        esum += np.bincount(bins, weights=p * rw * fft_scale2, minlength=nbins)
        count += np.bincount(bins, weights=rw, minlength=nbins).astype(np.int64)
    return esum, count.astype(np.int32), nbins, r_max, k_nyq


def _band_to_phys(S, band):
    """Padded inverse of a compact (N, N/2) band (ts.py:2639-2681)."""
    N, M = S.N, S.M
    Mh = M // 2 + 1
    H = N // 2
    tmp = np.zeros((M, Mh), dtype=S.cdtype)
    tmp[:N, :H] = band
    hi_end = min(3 * N // 4, Mh - 1)
    tmp[:, H:hi_end + 1] = 0
    tmp[N:N + H, :H] = tmp[H:N, :H]
    tmp[H:N, :H] = 0
    tmp[N, :H] = 0
    phys = spfft.irfft2(tmp, s=(M, M), axes=(0, 1), workers=S.workers, norm="forward")
    return phys.astype(S.dtype)


def kinetic(S):
    """sqrt(u^2+w^2) (ts.py:2629-2636)."""
    u, w = S.ur_full[0], S.ur_full[1]
    return np.sqrt(u * u + w * w).astype(S.dtype)


def om2_phys(S):
    """Physical vorticity (ts.py:2684-2687)."""
    return _band_to_phys(S, S.om2)


def stream_func(S):
    """Physical stream function from om2/(K2+1e-30) (ts.py:2690-2708)."""
    f = S.dtype
    K2 = (S.alfa[None, :] * S.alfa[None, :] + S.gamma[:, None] * S.gamma[:, None]).astype(f)
    K2 = (K2 + f(1.0e-30)).astype(f)
    return _band_to_phys(S, (S.om2 / K2).astype(S.cdtype))


def run(N, Re, K0, STEPS, CFL, spectrum="KM3", UPDATE=100, method="CNAB2",
        seed=1, dtype=np.float32, workers=None, fixed_dt=False, state=None):
    """run_dns loop (ts.py:2812-2848): returns (state, records).

    ``records`` holds (it, T, DT, CN, CFLM, TAU_L) at the NEXTDT cadence.
    ``fixed_dt`` skips NEXTDT (parity runs at fixed dt).
    """
    S = state or create_state(N, Re, K0, CFL, seed, spectrum, dtype, workers)
    init_dt(S)
    recs = []
    step = rk3 if method == "LS_IMEX_RK3" else cnab2
    for it in range(1, STEPS + 1):
        S.it = it
        dt_old = S.dt
        step(S)
        if not fixed_dt and ((it % UPDATE) == 0 or it == 1 or it == STEPS):
            CFLM = next_dt(S)
            _, _, tau = flow_eddy_metrics(S)
            recs.append((it, S.t, S.dt, S.cn, CFLM, tau))
        S.t += dt_old
    return S, recs
