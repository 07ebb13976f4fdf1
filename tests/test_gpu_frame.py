"""GPU tests of the display frame (palx_frame): pixels, statistics, metrics,
display stretch and scaling, the deferred NEXTDT and the viscosity controller,
each against the frame oracle (oracle/frame_oracle.py) or the handle's own
diagnostics, through the C ABI."""
import math

import numpy as np
import pytest

from oracle import frame_oracle as FO
from palinstrophy_amd import _palx as P
from palinstrophy_amd import frame as F

pytestmark = pytest.mark.gpu


def make(N, precision, K0=10.0, Re=1e4, seed=4895, steps=5, spectrum="KM3"):
    uc0, st = P.pao_generate(N, K0, Re, seed, spectrum == "KM3")
    h = P.Handle(N, precision)
    h.load_initial(uc0, P.calcom_f32(uc0))
    h.set_scalar(P.S_VISC, float(st[4]))
    h.set_scalar(P.S_CFLNUM, 0.5)
    h.set_scalar(P.S_DT, 0.5 / (h.cflm() * math.pi))
    if steps:
        h.step(steps, P.LS_IMEX_RK3, 0)
    return h


def run_frame(h, var, scale=1, flags=0, ctl=None):
    import torch
    w, hg, nb = h.frame_dims(scale)
    buf = torch.zeros(nb, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    h.frame(var, scale, flags, ctl, buf.data_ptr(), nb)
    h.sync()
    raw = buf.cpu().numpy()
    rec = raw[:P.FRAME_NREC * 8].view(np.float64).copy()
    pix = raw[P.FRAME_NREC * 8:].reshape(hg, w).copy()
    return rec, pix


def plane_of(h, var):
    return h.download(P.F_UR)[{P.VAR_U: 0, P.VAR_V: 1}.get(var, 2)]


@pytest.mark.parametrize("precision", [4, 8])
@pytest.mark.parametrize("N", [64, 256])
def test_pixels_and_statistics_match_oracle(N, precision):
    """[1,255] pixels bit-exact vs the oracle on the same plane, for every
    variable; min / max exact; mean / std within 1e-12."""
    print(f"case: frame pixels N={N} precision={precision}", flush=True)
    h = make(N, precision)
    for var in range(5):
        rec, pix = run_frame(h, var)
        plane = plane_of(h, var)
        ref = FO.float_to_pixels(plane)
        assert pix.shape == (h.M, h.M)
        assert np.array_equal(pix, ref), f"var {var}: {np.count_nonzero(pix != ref)} pixels differ"
        assert rec[P.FR_FMIN] == float(plane.min()) and rec[P.FR_FMAX] == float(plane.max())
        mu, sig = FO.pixel_stats(ref)
        assert abs(rec[P.FR_MU] - mu) <= 1e-12 * mu
        assert abs(rec[P.FR_SIGMA] - sig) <= 1e-12 * sig
        assert rec[P.FR_STOP] == 0.0 and rec[P.FR_WIDTH] == h.M and rec[P.FR_HEIGHT] == h.M
    h.close()


@pytest.mark.parametrize("precision", [4, 8])
def test_record_metrics_match_diag(precision):
    """U, L, tau_L, E(J), enstrophy, palinstrophy, kmax2 and the ratio of the
    record equal the blocking diagnostics of the same state."""
    h = make(128, precision)
    rec, _ = run_frame(h, P.VAR_OMEGA)
    d = h.diag()  # U, L, tau, Z, P, kmax2, E_J, E
    pairs = [(P.FR_U, d[0]), (P.FR_L, d[1]), (P.FR_TAU_L, d[2]), (P.FR_ENSTROPHY, d[3]),
             (P.FR_PALINSTROPHY, d[4]), (P.FR_KMAX2, d[5]), (P.FR_EJ, d[6]), (P.FR_ENERGY, d[7]),
             (P.FR_RATIO, d[4] / (d[3] * d[5]))]
    for idx, ref in pairs:
        assert abs(rec[idx] - ref) <= 1e-12 * abs(ref), (idx, rec[idx], ref)
    sc = h.scalars()
    assert rec[P.FR_DT] == sc[P.S_DT] and rec[P.FR_VISC] == sc[P.S_VISC] and rec[P.FR_CFLM] == sc[P.S_CFLM]
    h.close()


@pytest.mark.parametrize("scale", [1, 2, 5, -2, -3])
def test_stretch_and_display_scaling(scale):
    """Stretched display pixels at each scale: bit-exact vs the oracle's
    stretch of the [1,255] plane followed by its down / up scaling."""
    h = make(256, 4)
    rec, pix = run_frame(h, P.VAR_OMEGA, scale, P.FRAME_STRETCH)
    raw = FO.float_to_pixels(plane_of(h, P.VAR_OMEGA))
    sc = float(scale) if scale > 1 else (1.0 / -scale if scale < -1 else 1.0)
    ref = FO.upscale_downscale(FO.display_stretch(raw, rec[P.FR_MU], rec[P.FR_SIGMA]), sc)
    assert pix.shape == ref.shape
    assert np.array_equal(pix, ref)
    h.close()


def test_display_dims_follow_the_gui_scale():
    for N, expect in ((256, 768), (512, 768), (1024, 768), (4096, 1024)):
        h = P.Handle(N, 4)
        w, hg, nb = h.frame_dims(F.scale_arg(F.display_scale(N)))
        assert (w, hg) == (expect, expect) and nb == P.FRAME_NREC * 8 + w * hg
        h.close()


def test_nextdt_in_the_frame_equals_next_dt():
    a, b = make(128, 8), make(128, 8)
    rec, _ = run_frame(a, P.VAR_U, flags=P.FRAME_NEXTDT)
    b.next_dt()
    sa, sb = a.scalars(), b.scalars()
    assert sa[P.S_DT] == sb[P.S_DT] and sa[P.S_CN] == sb[P.S_CN]
    assert rec[P.FR_DT] == sb[P.S_DT] and rec[P.FR_CN] == sb[P.S_CN]
    a.close()
    b.close()


@pytest.mark.parametrize("precision", [4, 8])
def test_viscosity_controller_trajectory(precision):
    """Eight frames with the controller on: the device viscosity, Re and
    filter state follow the oracle controller fed with the device ratio, and
    a second run that sets the same viscosities from the host ends with a
    bit-identical omega (the steps read the device-written viscosity)."""
    N, K0 = 128, 10.0
    h = make(N, precision, K0=K0, steps=0)
    ref_run = make(N, precision, K0=K0, steps=0)
    rec0, _ = run_frame(h, P.VAR_OMEGA)
    # a set point away from the current state, so that the controller acts
    target = 0.5 * 10000 * rec0[P.FR_RATIO]
    ctl = F.frame_ctl(N, K0, target=target)
    h.frame_reset(1.0 / h.scalars()[P.S_VISC])
    oc = FO.ViscController(N, K0, target=target)
    visc = h.scalars()[P.S_VISC]
    changed = 0
    for f in range(8):
        h.step(4, P.LS_IMEX_RK3, 0)
        ref_run.step(4, P.LS_IMEX_RK3, 0)
        rec, _ = run_frame(h, P.VAR_OMEGA, 1, P.FRAME_ADAPT | P.FRAME_STRETCH, ctl)
        out = oc.adapt(rec[P.FR_RATIO], visc)
        assert abs(rec[P.FR_FILT] - oc._palin_filt) <= 1e-14 * abs(oc._palin_filt)
        if out is None:
            assert rec[P.FR_ADAPTED] == 0.0 and rec[P.FR_VISC] == visc
        else:
            changed += 1
            assert rec[P.FR_ADAPTED] == 1.0
            assert abs(rec[P.FR_VISC] - out[0]) <= 1e-14 * out[0]
            assert abs(rec[P.FR_RE] - out[1]) <= 1e-14 * out[1]
            visc = rec[P.FR_VISC]
        ref_run.set_scalar(P.S_VISC, rec[P.FR_VISC])
        print(f"  frame {f}: ratio {rec[P.FR_RATIO]:.6e} filt {rec[P.FR_FILT]:.4f} visc {rec[P.FR_VISC]:.6e}",
              flush=True)
    assert changed >= 4
    h.step(4, P.LS_IMEX_RK3, 0)
    ref_run.step(4, P.LS_IMEX_RK3, 0)
    assert np.array_equal(h.download(P.F_OM), ref_run.download(P.F_OM))
    h.close()
    ref_run.close()


def test_flat_field_stops_and_leaves_the_controller_alone():
    """A zero field maps to mid grey: sigma 0 -> stop, no controller update
    (turbo_main.py:2085-2087)."""
    N = 64
    h = P.Handle(N, 4)
    h.load_initial(np.zeros((2, N, N // 2), np.complex64), np.zeros((N, N // 2), np.complex64))
    h.set_scalar(P.S_VISC, 1e-3)
    h.frame_reset(1000.0)
    rec, pix = run_frame(h, P.VAR_OMEGA, 1, P.FRAME_ADAPT | P.FRAME_STRETCH, F.frame_ctl(N, 5.0))
    assert rec[P.FR_STOP] == 1.0 and rec[P.FR_SIGMA] == 0.0 and rec[P.FR_MU] == 128.0
    assert rec[P.FR_ADAPTED] == 0.0 and rec[P.FR_FILT] == 0.0 and rec[P.FR_VISC] == 1e-3
    assert rec[P.FR_RE] == 1000.0
    h.close()


def test_frame_argument_errors():
    h = make(64, 4, steps=1)
    import torch
    _, _, nb = h.frame_dims(1)
    buf = torch.zeros(nb, dtype=torch.uint8, device="cuda")
    with pytest.raises(ValueError):
        h.frame(7, 1, 0, None, buf.data_ptr(), nb)       # unknown variable
    with pytest.raises(ValueError):
        h.frame(0, 1, 64, None, buf.data_ptr(), nb)      # unknown flag
    with pytest.raises(ValueError):
        h.frame(0, 1, 0, None, buf.data_ptr(), nb - 1)   # buffer one byte short
    with pytest.raises(ValueError):
        h.frame(0, 1, P.FRAME_ADAPT, None, buf.data_ptr(), nb)  # controller without constants
    with pytest.raises(ValueError):
        h.frame_dims(h.M + 1)                            # downscale beyond the plane
    with pytest.raises(ValueError):
        h.frame_dims(-1000)                              # upscale beyond 16384 pixels
    h.stage_products()
    with pytest.raises(P.PalxError, match="products"):
        h.frame(0, 1, 0, None, buf.data_ptr(), nb)       # products in the work planes (state error)
    h.stage_update(P.LS_IMEX_RK3, 0)
    h.close()


def test_wrapper_frames_and_loop():
    """DnsSimulator.get_frame_pixels through palx_frame equals the oracle map
    of the plane; FrameLoop renders display-sized frames and mirrors the
    device scalars; a NaN field raises like sync_time_scalars_from_device."""
    from palinstrophy_amd.turbo_wrapper import DnsSimulator
    print("case: DnsSimulator + FrameLoop N=256", flush=True)
    sim = DnsSimulator(n=256, re=1000.0, k0=10.0, cfl=0.5)
    for _ in range(3):
        sim.step(2)
    for var in range(5):
        sim.set_variable(var)
        pix = sim.get_frame_pixels()
        ref = FO.float_to_pixels(sim.state.handle.download(P.F_UR)[{0: 0, 1: 1}.get(var, 2)])
        assert np.array_equal(pix, ref)
    assert sim.dt == sim.state.handle.scalars()[P.S_DT]
    loop = F.FrameLoop(sim, update_interval=3, adapt=True, target=5.0)
    frames = [loop.tick() for _ in range(9)]
    shown = [f for f in frames if f is not None]
    assert len(shown) == 3 and all(f.pixels.shape == (768, 768) for f in shown)
    assert len(loop.csv_rows) == 3 and len(loop.csv_rows[0]) == len(F.CSV_HEADER)
    assert sim.state.visc == sim.state.handle.scalars()[P.S_VISC]
    # set_N replaces the state: the loop rebinds (N=128 -> 192^2 padded, shown 5x = 960^2)
    sim.set_N(128)
    assert loop.render().pixels.shape == (960, 960)
    sim.set_N(256)
    uc0 = np.zeros((2, 256, 128), np.complex64)
    uc0[0, 3, 2] = np.nan
    sim.state.handle.load_initial(uc0, np.zeros((256, 128), np.complex64))
    with pytest.raises(FloatingPointError):
        sim.get_frame_pixels()
