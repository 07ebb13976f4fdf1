"""GPU tests beyond the per-step parity: the N=1024 PAO RK3 and N=256 KM3 CNAB2 workloads,
the FP32 bar, the reference-compatible facade end to end, getters,
diagnostics, spectrum, full-size (N=4096) checks and failure handling."""
import json
import math
import os

import numpy as np
import pytest

from oracle import dns_oracle as O
from palinstrophy_amd import _palx as P

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rel_l2(a, b):
    a = np.asarray(a).astype(np.complex128).ravel()
    b = np.asarray(b).astype(np.complex128).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def pair(N, Re, K0, CFL, seed, spectrum, precision, oracle_dtype=np.float64):
    S = O.create_state(N, Re, K0, CFL, seed, spectrum, dtype=oracle_dtype)
    uc0, _ = P.pao_generate(N, K0, Re, seed, spectrum == "KM3")
    assert np.array_equal(uc0, S.uc0)
    h = P.Handle(N, precision)
    h.load_initial(uc0, P.calcom_f32(uc0))
    h.set_scalar(P.S_VISC, S.visc)
    h.set_scalar(P.S_CFLNUM, CFL)
    return S, h


def fields(h):
    om = h.download(P.F_OM)
    h.physical()
    ur = h.download(P.F_UR)
    return om, ur


@pytest.mark.timeout(600)
def test_pao_N1024_rk3_fp64_100_fixed_steps():
    """N=1024, Re=1e4, PAO K0=15, seed 4895, LS-IMEX-RK3, dt = 0.1/(CFLM0 pi),
    100 fixed-dt steps: omega, u, v relative L2 <= 1e-10 vs the FP64 oracle."""
    print("case: PAO N=1024 RK3 fp64 100 fixed-dt steps", flush=True)
    S, h = pair(1024, 1e4, 15.0, 0.1, 4895, "PAO", 8)
    O.init_dt(S)
    h.set_scalar(P.S_DT, S.dt)
    h.step(100, P.LS_IMEX_RK3, 0)
    for _ in range(100):
        O.rk3(S)
    om, ur = fields(h)
    e = (rel_l2(om, S.om2), rel_l2(ur[0], S.ur_full[0]), rel_l2(ur[1], S.ur_full[1]))
    print(f"  rel L2 omega={e[0]:.2e} u={e[1]:.2e} v={e[2]:.2e}", flush=True)
    assert max(e) <= 1e-10


@pytest.mark.timeout(600)
def test_km3_N256_cnab2_fp64_adaptive():
    """sim 256 10000 10 1001 0.75 KM3 (seed 4895), CNAB2 with NEXTDT at
    it==1 / it%100==0 / it==STEPS.  The case diverges in the reference
    algorithm between steps 200 and 300 (CFL 0.75 is beyond CNAB2's
    stability with a 100-step NEXTDT cadence; the oracle shows it in FP32
    and FP64).  Parity: omega, u, v, dt, t, E, Z at step 200 within 1e-10 of
    the FP64 oracle, and the same non-finite end state at step 1001."""
    print("case: KM3 N=256 CNAB2 fp64 adaptive", flush=True)
    S, h = pair(256, 1e4, 10.0, 0.75, 4895, "KM3", 8)
    c0 = O.init_dt(S)
    g0 = h.cflm()
    assert abs(g0 - c0) < 1e-12 * c0
    h.set_scalar(P.S_DT, S.dt)
    h.set_iteration(0)
    h.step(200, P.CNAB2, 100)            # NEXTDT after it = 1, 100, 200
    S2, recs = O.run(256, 1e4, 10.0, 200, 0.75, "KM3", 100, "CNAB2", seed=4895, dtype=np.float64)
    om, ur = fields(h)
    sc = h.scalars()
    e_om, e_u = rel_l2(om, S2.om2), rel_l2(ur[0:2], S2.ur_full[0:2])
    print(f"  step 200: rel L2 omega={e_om:.2e} u,v={e_u:.2e} dt {sc[P.S_DT]:.15e} vs {S2.dt:.15e}", flush=True)
    assert abs(sc[P.S_DT] - S2.dt) <= 1e-12 * S2.dt
    assert abs(sc[P.S_T] - S2.t) <= 1e-12 * S2.t
    assert e_om <= 1e-10 and e_u <= 1e-10
    d = h.diag()
    pe, Zo, Po, _ = O.pal_over_ens_kmax2(S2)
    assert abs(d[3] - Zo) <= 1e-10 * Zo and abs(d[4] - Po) <= 1e-10 * Po
    assert abs(d[6] - O.energy_joules(S2)) <= 1e-10 * d[6]
    # continue to STEPS=1001 like `sim`: both blow up
    h.step(800, P.CNAB2, 100)
    h.step(1, P.CNAB2, 0)
    h.next_dt()
    S3, _ = O.run(256, 1e4, 10.0, 1001, 0.75, "KM3", 100, "CNAB2", seed=4895, dtype=np.float64)
    sc = h.scalars()
    assert not np.isfinite(S3.dt) and not np.isfinite(sc[P.S_DT])
    assert sc[P.S_NAN] == 1.0


@pytest.mark.parametrize("N,method", [(256, "LS_IMEX_RK3"), (512, "CNAB2")])
def test_fp32_against_fp32_and_fp64_oracles(N, method):
    """FP32 bulk path: tight against the FP32 oracle (same algorithm, FFT
    rounding only) and within 1e-4 of the FP64 oracle after 100 steps."""
    print(f"case: fp32 N={N} {method} 100 steps", flush=True)
    S64, h = pair(N, 1e4, 10.0, 0.5, 4895, "KM3", 4)
    S32 = O.create_state(N, 1e4, 10.0, 0.5, 4895, "KM3", dtype=np.float32)
    O.init_dt(S64)
    O.init_dt(S32)
    S32.dt = S64.dt
    h.set_scalar(P.S_DT, S64.dt)
    m = P.CNAB2 if method == "CNAB2" else P.LS_IMEX_RK3
    step = O.cnab2 if method == "CNAB2" else O.rk3
    for _ in range(100):
        step(S64)
        step(S32)
    h.step(100, m, 0)
    om, ur = fields(h)
    e64 = rel_l2(om, S64.om2)
    e32 = rel_l2(om, S32.om2)
    ref32 = rel_l2(S32.om2, S64.om2)
    print(f"  fp32 GPU vs fp64 oracle {e64:.2e}; vs fp32 oracle {e32:.2e}; fp32 oracle vs fp64 {ref32:.2e}")
    assert e64 <= 1e-4
    assert e64 <= 3.0 * ref32 + 1e-6


def test_getters_diag_spectrum_fp64():
    print("case: getters/diag/spectrum N=64 fp64", flush=True)
    S, h = pair(64, 1e4, 10.0, 0.5, 4895, "KM3", 8)
    O.init_dt(S)
    h.set_scalar(P.S_DT, S.dt)
    for _ in range(5):
        O.rk3(S)
    h.step(5, P.LS_IMEX_RK3, 0)
    h.getter(P.VAR_ENERGY)
    ur = h.download(P.F_UR)
    assert rel_l2(ur[2], O.kinetic(S)) < 1e-12
    h.getter(P.VAR_OMEGA)
    assert rel_l2(h.download(P.F_UR)[2], O.om2_phys(S)) < 1e-12
    h.getter(P.VAR_STREAM)
    assert rel_l2(h.download(P.F_UR)[2], O.stream_func(S)) < 1e-12
    d = h.diag()
    U, L, tau = O.flow_eddy_metrics(S)
    assert abs(d[0] - U) < 1e-12 * U and abs(d[1] - L) < 1e-12 * L and abs(d[2] - tau) < 1e-12 * tau
    pe, Z, Pal, kmax2 = O.pal_over_ens_kmax2(S)
    assert abs(d[3] - Z) < 1e-12 * Z and abs(d[4] - Pal) < 1e-12 * Pal and d[5] == kmax2
    e, c, nbins, *_ = O.spectrum_bins(S)
    ge, gc = h.spectrum(nbins)
    assert np.array_equal(gc, c)
    assert np.allclose(ge, e, rtol=1e-12, atol=0)


def test_facade_run_dns_matches_reference_log(capsys):
    """The reference CLI case `sim 512 10000 10 1001 0.25 gpu PAO` through the
    reference-named API: every ITERATION record within the reference's own
    CPU/GPU spread of its log."""
    from palinstrophy_amd import turbo_simulator as ts
    print("case: facade run_dns N=512 CNAB2 fp32 1001 steps", flush=True)
    os.environ.pop("SCIPYTURBO_SEED", None)
    ts.run_dns(512, 1e4, 10.0, 1001, 0.25, "gpu", "PAO", 100, "CNAB2", precision="fp32")
    out = capsys.readouterr().out
    recs = [l.split() for l in out.splitlines() if l.strip().startswith("ITERATION")]
    ref = json.load(open(os.path.join(GOLD, "log_sim_512_1001.json")))["cpu_linux"]["iterations"]
    assert len(recs) == len(ref)
    for r, q in zip(recs, ref):
        it = int(r[1])
        vals = {k: float(v) for k, v in (x.split("=") for x in r[2:])}
        assert it == q[0]
        assert abs(vals["DT"] - float(q[2])) <= 1e-5 * float(q[2]) + 5e-9
        assert abs(vals["CFLM"] - float(q[4])) <= 2e-5 * float(q[4])
        assert abs(vals["TAU_L"] - float(q[5])) <= 2e-6 * float(q[5]) + 1e-6


def test_facade_wrapper_and_getters():
    from palinstrophy_amd.turbo_wrapper import DnsSimulator
    print("case: DnsSimulator N=64", flush=True)
    sim = DnsSimulator(n=64, re=1000.0, k0=5.0, cfl=0.5)
    for i in range(4):
        sim.step(2)
    for var in range(5):
        sim.set_variable(var)
        pix = sim.get_frame_pixels()
        assert pix.shape == (96, 96) and pix.dtype == np.uint8
        assert pix.min() >= 1
    assert sim.get_iteration() == 4 and sim.get_time() > 0


@pytest.mark.timeout(600)
def test_full_size_N4096_fp64_two_rk3_steps_vs_oracle():
    """Full headline size: 2 fixed-dt RK3 steps of N=4096 (6144^2) in FP64
    against the FP64 oracle."""
    print("case: N=4096 fp64 2 RK3 steps vs oracle", flush=True)
    N, K0 = 4096, 15.0
    Re = O.re_from_n_k0(N, K0)
    uc0, st = P.pao_generate(N, K0, Re, 4895, True)
    S = O.create_state(N, Re, K0, 0.5, 4895, "KM3", dtype=np.float64, uc0=uc0)
    h = P.Handle(N, 8)
    h.load_initial(uc0, P.calcom_f32(uc0))
    h.set_scalar(P.S_VISC, S.visc)
    O.init_dt(S)
    h.set_scalar(P.S_DT, S.dt)
    for _ in range(2):
        O.rk3(S)
    h.step(2, P.LS_IMEX_RK3, 0)
    om = h.download(P.F_OM)
    e = rel_l2(om, S.om2)
    print(f"  rel L2 omega={e:.2e}", flush=True)
    assert e <= 1e-10
    h.close()


@pytest.mark.timeout(600)
def test_full_size_fp32_path_against_fp64_path_100_steps():
    """Precision form at N=4096 (KM3, K0=15, Re from the Re(N,K0) rule, seed
    4895, RK3, CFL 0.5, each path with its own adaptive dt): after 100 steps
    the FP32 path's omega is within 1e-4 (relative L2) of the FP64 path's,
    its energy spectrum within 1e-4 and its enstrophy within 1e-5."""
    print("case: N=4096 fp32 vs fp64 GPU, 100 adaptive RK3 steps", flush=True)
    N, K0 = 4096, 15.0
    Re = O.re_from_n_k0(N, K0)
    uc0, st = P.pao_generate(N, K0, Re, 4895, True)
    om0 = P.calcom_f32(uc0)
    hs = {}
    for prec in (4, 8):
        h = P.Handle(N, prec)
        h.load_initial(uc0, om0)
        h.set_scalar(P.S_VISC, st[4])
        h.set_scalar(P.S_CFLNUM, 0.5)
        h.set_scalar(P.S_DT, 0.5 / (h.cflm() * math.pi))
        h.step(100, P.LS_IMEX_RK3, 100)
        hs[prec] = h
    om32 = hs[4].download(P.F_OM)
    om64 = hs[8].download(P.F_OM)
    e_om = rel_l2(om32, om64)
    s32, _ = hs[4].spectrum(2 * N)
    s64, _ = hs[8].spectrum(2 * N)
    e_sp = float(np.linalg.norm(s32 - s64) / np.linalg.norm(s64))
    z32, z64 = hs[4].diag()[3], hs[8].diag()[3]
    e_z = abs(z32 - z64) / abs(z64)
    print(f"  omega {e_om:.2e}  E(k) {e_sp:.2e}  enstrophy {e_z:.2e}", flush=True)
    assert e_om <= 1e-4 and e_sp <= 1e-4 and e_z <= 1e-5
    for h in hs.values():
        h.close()


@pytest.mark.parametrize("N,prec,tol", [(4096, 8, 1e-12), (4096, 4, 2e-5), (6144, 8, 1e-12), (6144, 4, 2e-5),
                                        (8192, 8, 1e-12), (8192, 4, 2e-5), (12288, 8, 1e-12), (12288, 4, 2e-5),
                                        (16384, 4, 2e-5)])
def test_full_size_single_mode_closed_form(N, prec, tol):
    """N=4096 and the largest grid, N=6144 (9216^2 padded): a single Fourier
    mode has no nonlinear term; after K RK3 steps its amplitude is the
    closed-form product of the stage factors and every other mode stays at
    roundoff.  N=16384 is an FP32-only size; N=16384, and N=8192 and 12288
    in FP64, have the split-exchange variants only."""
    print(f"case: single mode N={N} precision={prec}", flush=True)
    H = N // 2
    kz, kx = 37, 211
    om0 = 0.7 - 0.4j
    k2 = kx * kx + kz * kz
    cd = np.complex128 if prec == 8 else np.complex64
    om = np.zeros((N, H), dtype=cd)
    om[kz, kx] = om0
    uc = np.zeros((2, N, H), dtype=cd)
    uc[0, kz, kx] = -1j * kz * om0 / k2
    uc[1, kz, kx] = 1j * kx * om0 / k2
    h = P.Handle(N, prec)
    h.upload(P.F_OM, om)
    h.load_velocity(uc)
    nu, dt, K = 1e-5, 1e-3, 5
    h.set_scalar(P.S_VISC, nu)
    h.set_scalar(P.S_DT, dt)
    h.step(K, P.LS_IMEX_RK3, 0)
    g = 1.0
    for al, be in zip(O.RK3_ALPHA, O.RK3_BETA):
        g *= (1 - k2 * be * dt * nu) / (1 + k2 * al * dt * nu)
    w = h.download(P.F_OM)
    expect = complex(om[kz, kx]) * g ** K
    assert abs(w[kz, kx] - expect) <= tol * abs(expect)
    w[kz, kx] = 0
    assert np.abs(w).max() <= tol * abs(expect)
    h.close()


@pytest.mark.parametrize("prec,tol", [(8, 1e-12), (4, 2e-5)])
def test_n6144_every_compiled_variant_single_mode(prec, tol):
    """N=6144 (9216-point transforms) has the base plan in variants 0-2 and the
    split exchange in variant 6: every compiled one, for both passes, keeps a
    single mode on its closed-form decay with the other modes at roundoff."""
    N, H = 6144, 3072
    kz, kx = 37, 211
    om0 = 0.7 - 0.4j
    k2 = kx * kx + kz * kz
    cd = np.complex128 if prec == 8 else np.complex64
    om = np.zeros((N, H), dtype=cd)
    om[kz, kx] = om0
    uc = np.zeros((2, N, H), dtype=cd)
    uc[0, kz, kx] = -1j * kz * om0 / k2
    uc[1, kz, kx] = 1j * kx * om0 / k2
    nu, dt, K = 1e-5, 1e-3, 3
    g = 1.0
    for al, be in zip(O.RK3_ALPHA, O.RK3_BETA):
        g *= (1 - k2 * be * dt * nu) / (1 + k2 * al * dt * nu)
    expect = complex(om0) * g ** K
    vs = [v for v in range(P.n_variants(N)) if P.variant_ok(N, prec, v)]
    assert 6 in vs and 3 not in vs
    for v in vs:
        print(f"case: single mode N={N} precision={prec} variant {v}", flush=True)
        h = P.Handle(N, prec)
        h.set_variants(v, v)
        h.upload(P.F_OM, om)
        h.load_velocity(uc)
        h.set_scalar(P.S_VISC, nu)
        h.set_scalar(P.S_DT, dt)
        h.step(K, P.LS_IMEX_RK3, 0)
        w = h.download(P.F_OM)
        assert abs(w[kz, kx] - expect) <= tol * abs(expect), v
        w[kz, kx] = 0
        assert np.abs(w).max() <= tol * abs(expect), v
        h.close()


def test_nan_is_detected_and_zero_field_keeps_dt():
    from palinstrophy_amd import turbo_simulator as ts
    S = ts.create_dns_state(32, 1e3, 3.0, 0.5, "gpu", 5, start_spectrum="PAO", precision="fp64")
    # zero field: CFLM = 0 -> dt unchanged (turbo_simulator.py:2161-2162)
    S.om2 = np.zeros((32, 16), dtype=np.complex128)
    S.handle.load_initial(np.zeros((2, 32, 16), np.complex64), np.zeros((32, 16), np.complex64))
    S.dt = 1e-3
    assert ts.next_dt(S, sync_host=True) == 0.0
    assert S.dt == 1e-3
    # a NaN planted from the host in the spectral state surfaces as FloatingPointError
    uc0 = np.zeros((2, 32, 16), np.complex64)
    uc0[0, 3, 2] = np.nan
    S.handle.load_initial(uc0, np.zeros((32, 16), np.complex64))
    with pytest.raises(FloatingPointError):
        ts.next_dt(S, sync_host=True)


def test_facade_arrays_are_live_zero_copy_views():
    """S.om2 / S.fnm1 / S.ur_full alias the handle's device buffers (the
    reference's CuPy arrays are live arrays too): no copy on access, they
    follow the steps, and assigning a CUDA tensor copies on the device."""
    import torch
    from palinstrophy_amd import turbo_simulator as ts

    print("case: facade zero-copy views N=64 fp64", flush=True)
    S = ts.create_dns_state(N=64, Re=1e4, K0=10.0, CFL=0.5, backend="gpu", seed=7, start_spectrum="PAO",
                            precision="fp64")
    om = S.om2
    assert om.is_cuda and tuple(om.shape) == (64, 32) and om.dtype == torch.complex128
    assert om.data_ptr() == S.handle.device_ptr(P.F_OM)[0]
    assert np.array_equal(om.cpu().numpy(), S.handle.download(P.F_OM))
    ur = S.ur_full
    assert tuple(ur.shape) == (3, 96, 96) and ur.data_ptr() == S.handle.device_ptr(P.F_UR)[0]
    S.dt = 0.5 / (ts.compute_cflm(S) * math.pi)
    ts.dns_step_ls_imex_rk3(S)
    torch.cuda.synchronize()
    assert np.array_equal(om.cpu().numpy(), S.handle.download(P.F_OM))  # the old view follows the step
    assert np.array_equal(S.fnm1.cpu().numpy(), S.handle.download(P.F_FNM1))
    x = om.clone() * 2
    S.om2 = x
    torch.cuda.synchronize()
    assert torch.equal(S.om2, x)
    assert np.array_equal(S.handle.download(P.F_OM), x.cpu().numpy())


def test_spectrum_bins_are_bit_reproducible_and_match_oracle():
    """Each bin is summed by one wavefront in a fixed order: repeated calls
    (host and device outputs) give the same bits; the bins follow the FP64
    oracle's on the same omega (N=1024 FP32 after 20 RK3 steps, N=3072 FP64
    on its seeded field, every bin of a 3*2^k grid)."""
    for N, prec, steps in ((1024, 4, 20), (3072, 8, 0)):
        print(f"case: spectrum reproducibility N={N} precision={prec}", flush=True)
        S, h = pair(N, 1e4, 15.0, 0.5, 4895, "KM3", prec)
        if steps:
            h.set_scalar(P.S_DT, 1e-4)
            h.step(steps, P.LS_IMEX_RK3, 0)
        nbins = 2 * N
        e1, c1 = h.spectrum(nbins)
        e2, c2 = h.spectrum(nbins)
        import torch
        de = torch.empty(nbins, dtype=torch.float64, device="cuda")
        dc = torch.empty(nbins, dtype=torch.int32, device="cuda")
        h.spectrum_into(de.data_ptr(), dc.data_ptr(), nbins)
        h.sync()
        assert np.array_equal(e1, e2) and np.array_equal(c1, c2)
        assert np.array_equal(de.cpu().numpy(), e1) and np.array_equal(dc.cpu().numpy(), c1)
        om = h.download(P.F_OM).astype(np.complex128)
        S.arrays["om2"] = om
        eo, co, *_ = O.spectrum_bins(S)
        assert np.array_equal(c1, co)
        assert np.abs(e1 - eo).max() <= 1e-12 * eo.max()
        h.close()


def test_spectrum_time_average_on_device_matches_host(tmp_path):
    """The trapezoid-in-time spectrum average accumulated on the device (no
    host transfer per sample) equals the host accumulation of the same
    samples (turbo_simulator.py:2386-2497), and writes the same CSV."""
    from palinstrophy_amd import turbo_simulator as ts

    print("case: device spectrum average N=64 fp64", flush=True)
    S = ts.create_dns_state(N=64, Re=1e4, K0=10.0, CFL=0.5, backend="gpu", seed=11, start_spectrum="KM3",
                            precision="fp64")
    S.dt = 0.5 / (ts.compute_cflm(S) * math.pi)
    host, dev = ts.SpectrumAverageState(), ts.SpectrumAverageState()
    t = 0.0
    for it in range(6):
        assert ts.sample_and_update_spectrum_average(S, host, t, it)
        assert ts.sample_and_update_spectrum_average(S, dev, t, it, on_device=True)
        assert dev.integral_energy.is_cuda  # stays on the device
        ts.dns_step_ls_imex_rk3(S)
        t += S.dt
    assert ts.sample_and_update_spectrum_average(S, dev, t, 6)  # keeps living on the device
    assert ts.sample_and_update_spectrum_average(S, host, t, 6)
    assert dev.sample_count == host.sample_count == 7 and dev.total_dt == host.total_dt
    mh, md = ts.spectrum_average_mean_energy(host), ts.spectrum_average_mean_energy(dev)
    assert np.allclose(md, mh, rtol=1e-12, atol=0)
    assert np.array_equal(dev.count.cpu().numpy(), host.count)
    ts.save_spectrum_average_csv(dev, str(tmp_path / "d.csv"), ["N"], [64])
    ts.save_spectrum_average_csv(host, str(tmp_path / "h.csv"), ["N"], [64])
    rows_d = (tmp_path / "d.csv").read_text().splitlines()
    rows_h = (tmp_path / "h.csv").read_text().splitlines()
    assert len(rows_d) == len(rows_h) and rows_d[:3] == rows_h[:3]


@pytest.mark.timeout(600)
@pytest.mark.parametrize("n1,n2", [(4096, 8192), (6144, 12288), (4096, 16384)])
def test_fp32_only_grid_matches_half_grid_on_shared_modes(n1, n2):
    """N=8192 (12288^2 padded), N=12288 (18432^2) and N=16384 (24576^2)
    against a grid of half or a quarter the size (N=4096, N=6144) on the same
    field, in FP32, and in FP64 where the large grid has it (8192, 12288:
    split-exchange variants; the FP64 path must then match to 1e-12 / 1e-10).  A PAO field with K0=15 is
    band-limited far below k=n1/2, so
      * its physical u, w on the n2 grid, sampled at every (n2/n1)-th point,
        are its values on the n1 grid (both are exact trigonometric
        interpolations): within FP32 rounding, 1e-6;
      * its dealiased nonlinear term is exact on both grids: after one CNAB2
        step from fnm1 = 0, fnm1 holds fn.  Against the n1 FP64 fn, the n2
        FP32 fn on the shared modes must be as accurate as the n1 FP32 fn
        (within 2x: FP32 FFT noise, amplified by kz kx at high k, dominates
        both), and on the energetic modes |kx|, |kz| < 128 within 1e-5.
    Every kernel variant compiled for n2 passes."""
    print(f"case: N={n2} fp32 vs N={n1} (physical field, nonlinear term)", flush=True)
    if n2 == 16384:
        with pytest.raises(ValueError):
            P.Handle(n2, 8)
    uc_s, st = P.pao_generate(n1, 15.0, 1e5, 4895, False)
    om_s = P.calcom_f32(uc_s)
    z = np.arange(n1)
    kz = np.where(z < n1 // 2, z, z - n1)
    rows = kz % n2  # row of each N=4096 kz on the N=8192 grid
    low = (np.abs(kz)[:, None] < 128) & (np.arange(n1 // 2)[None, :] < 128)

    def embed(a):
        out = np.zeros(a.shape[:-2] + (n2, n2 // 2), dtype=a.dtype)
        out[..., rows, : n1 // 2] = a
        return out

    def run(N, prec, uc0, om0, v):
        h = P.Handle(N, prec)
        h.set_variants(v, v)
        h.load_initial(uc0, om0)
        h.physical()
        u = h.download(P.F_UR)[0:2].astype(np.float64)
        h.set_scalar(P.S_VISC, float(st[4]))
        h.set_scalar(P.S_DT, 2e-4)
        h.step(1, P.CNAB2, 0)
        fn = h.download(P.F_FNM1).astype(np.complex128)
        h.close()
        return u, fn

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    _, fn64 = run(n1, 8, uc_s, om_s, 0)
    u1, fn1 = run(n1, 4, uc_s, om_s, 0)
    e1 = rel(fn1, fn64)
    uc_b, om_b = embed(uc_s), embed(om_s)
    step = n2 // n1
    for v in range(P.n_variants(n2)):
        if not P.variant_ok(n2, 4, v):
            continue
        u2, fn2 = run(n2, 4, uc_b, om_b, v)
        fs = fn2[rows, : n1 // 2]
        eu, ef, el = rel(u2[:, ::step, ::step], u1), rel(fs, fn64), rel(fs[low], fn64[low])
        print(f"  variant {v}: physical {eu:.2e}; fn vs FP64 {ef:.2e} (N={n1} FP32: {e1:.2e}), "
              f"low modes {el:.2e}", flush=True)
        assert eu <= 1e-6 and ef <= 2 * e1 and el <= 1e-5
    if n2 != 16384:  # the FP64 path at n2 (split-exchange variants) against FP64 at n1
        u64, _ = run(n1, 8, uc_s, om_s, 0)
        for v in range(P.n_variants(n2)):
            if not P.variant_ok(n2, 8, v):
                continue
            u2, fn2 = run(n2, 8, uc_b, om_b, v)
            fs = fn2[rows, : n1 // 2]
            eu, ef = rel(u2[:, ::step, ::step], u64), rel(fs, fn64)
            print(f"  FP64 variant {v}: physical {eu:.2e}; fn vs N={n1} FP64 {ef:.2e}", flush=True)
            assert eu <= 1e-12 and ef <= 1e-10


def test_pgm_dump_known_answers(tmp_path):
    """dump_field_as_pgm_full (turbo_simulator.py:2173-2222): header, size and
    the round-half-up [1,255] map: 0 -> 1, 0.25 -> 65 (64.5 + 0.5), 0.5 -> 128,
    1 -> 255; a constant plane -> 128."""
    import torch
    from palinstrophy_amd import turbo_simulator as ts
    S = ts.create_dns_state(8, 1e3, 3.0, 0.5, "gpu", 5, start_spectrum="PAO", precision="fp32")
    plane = torch.zeros((12, 12), dtype=torch.float32)
    plane[0, 1], plane[0, 2], plane[0, 3] = 0.25, 0.5, 1.0
    S.ur_full[2].copy_(plane.to("cuda"))
    path = tmp_path / "w.pgm"
    ts.dump_field_as_pgm_full(S, 2, str(path))
    raw = path.read_bytes()
    head = b"P5\n12 12\n255\n"
    assert raw[:len(head)] == head and len(raw) == len(head) + 144
    pix = np.frombuffer(raw[len(head):], dtype=np.uint8).reshape(12, 12)
    assert pix[0, :4].tolist() == [1, 65, 128, 255] and (pix[1:] == 1).all()
    S.ur_full[2].fill_(3.0)
    ts.dump_field_as_pgm_full(S, 2, str(path))
    assert (np.frombuffer(path.read_bytes()[len(head):], dtype=np.uint8) == 128).all()


@pytest.mark.parametrize("N,prec", [(8, 8), (8, 4), (32, 8), (32, 4), (48, 4), (256, 8), (4096, 4)])
def test_work_planes_hold_the_padded_z_inverse_tiled_by_z(N, prec):
    """Velocity work planes after palx_load_initial: the 3/2-padded, unscaled
    z-inverse of the compact spectra for kx < N/2 (turbo_simulator.py:2021-2034,
    1255-1262), stored in tiles of P elements along z (P = 4 FP64, 4 FP32 up
    to M = 3072 and 8 above, halved until it divides M)."""
    print(f"case: work-plane layout N={N} prec={prec}", flush=True)
    rng = np.random.default_rng(N + prec)
    H, M = N // 2, 3 * N // 2
    uc0 = (rng.standard_normal((2, N, H)) + 1j * rng.standard_normal((2, N, H))).astype(np.complex64)
    uc0[:, N // 2, :] = 0  # the kz = -N/2 row, dropped by the 3/2 rule
    h = P.Handle(N, prec)
    h.load_initial(uc0, P.calcom_f32(uc0))
    raw = h.download(P.F_WORK)
    h.close()
    assert raw.shape == (2, M // h.work_tile, H, h.work_tile)
    assert h.work_tile == (8 if prec == 4 and 3 * N // 2 > 3072 else 4)
    pad = np.zeros((2, M, H), np.complex128)
    pad[:, :N // 2] = uc0[:, :N // 2]
    pad[:, M - N // 2:] = uc0[:, N // 2:]
    ref = (np.fft.ifft(pad, axis=1) * M).transpose(0, 2, 1)  # [c][kx][z]
    err = np.abs(P.untile_work(raw) - ref).max() / np.abs(ref).max()
    print(f"  max rel err {err:.2e}", flush=True)
    assert err < (1e-13 if prec == 8 else 2e-6)


@pytest.mark.parametrize("prec,tol", [(8, 1e-14), (4, 1e-6)])
def test_full_size_nonlinear_term_conserves_energy_and_enstrophy(prec, tol):
    """N=4096 (6144^2 padded), a size-independent invariant of the dealiased
    nonlinear term (turbo_simulator.py:1520-1552): with the velocity
    reconstructed from omega (palx_recon_velocity), the term fn the column
    pass forms transfers no energy and no enstrophy, Re sum w conj(om) fn / k^2
    = Re sum w conj(om) fn = 0 to rounding (w = 1 at kx = 0, 2 otherwise),
    relative to sum w |om| |fn| (/ k^2).  PAO: the kz = -N/2 row is zero."""
    print(f"case: N=4096 Galerkin invariants prec={prec}", flush=True)
    N, H = 4096, 2048
    uc0, st = P.pao_generate(N, 15.0, 1e4, 4895, False)
    om0 = P.calcom_f32(uc0)
    h = P.Handle(N, prec)
    h.load_initial(uc0, om0)
    h.recon_velocity()
    h.set_scalar(P.S_VISC, st[4])
    h.set_scalar(P.S_DT, 1e-4)
    h.stage_products()
    h.stage_update(P.LS_IMEX_RK3, 0)
    fn = h.download(P.F_FNM1).astype(np.complex128)
    h.close()
    om = om0.astype(np.complex128)
    z = np.arange(N)
    kz = np.where(z < N // 2, z, z - N)[:, None].astype(float)
    kx = np.arange(H)[None, :].astype(float)
    k2 = kx * kx + kz * kz
    w = np.where(kx == 0, 1.0, 2.0)
    inv = np.where(k2 > 0, 1.0 / np.where(k2 > 0, k2, 1.0), 0.0)
    t = w * (np.conj(om) * fn).real
    a = w * np.abs(om) * np.abs(fn)
    dE, dZ = abs(np.sum(t * inv)) / np.sum(a * inv), abs(np.sum(t)) / np.sum(a)
    print(f"  energy transfer {dE:.2e}  enstrophy transfer {dZ:.2e} (relative)", flush=True)
    assert np.abs(fn).max() > 0
    assert dE < tol and dZ < tol
