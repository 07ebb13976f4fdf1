"""GPU parity: the HIP DNS step (libpalx via ctypes) against the CPU oracle.

FP64 bar: relative L2 <= 1e-10 on spectral omega and on the
physical u, w after fixed-dt steps from the same seeded initial field.
FP32 bar: relative L2 <= 1e-4 against the FP64 oracle after 100 steps.
"""
import math

import numpy as np
import pytest

from oracle import dns_oracle as O
from palinstrophy_amd import _palx as P

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a).astype(np.complex128).ravel()
    b = np.asarray(b).astype(np.complex128).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def start_pair(N, Re, K0, CFL, seed, spectrum, precision, oracle_dtype=np.float64):
    """Oracle state + GPU handle from the same complex64 seeded field."""
    S = O.create_state(N, Re, K0, CFL, seed, spectrum, dtype=oracle_dtype)
    h = P.Handle(N, precision)
    om0 = P.calcom_f32(S.uc0)
    h.load_initial(S.uc0, om0)
    h.set_scalar(P.S_VISC, S.visc)
    h.set_scalar(P.S_CFLNUM, CFL)
    return S, h


def test_initial_cflm_and_physical_fp64():
    print("case: initial physical field N=64 KM3 seed 4895 fp64", flush=True)
    S, h = start_pair(64, 1e4, 10, 0.75, 4895, "KM3", 8)
    cflm0 = O.init_dt(S)
    g = h.cflm()
    ur = h.download(P.F_UR)
    assert rel_l2(ur[0:2], S.ur_full[0:2]) < 1e-12
    assert abs(g - cflm0) <= 1e-12 * cflm0


@pytest.mark.parametrize("N,method,spectrum", [
    (16, "CNAB2", "PAO"), (32, "CNAB2", "KM3"), (64, "LS_IMEX_RK3", "KM3"),
    (128, "CNAB2", "PAO"), (256, "LS_IMEX_RK3", "PAO")])
def test_fixed_dt_steps_fp64(N, method, spectrum):
    nsteps = 20
    print(f"case: fixed-dt parity N={N} {method} {spectrum} fp64 steps={nsteps}", flush=True)
    S, h = start_pair(N, 1e4, 10 if N >= 32 else 3, 0.5, 4895, spectrum, 8)
    O.init_dt(S)
    h.set_scalar(P.S_DT, S.dt)
    h.set_scalar(P.S_CN, 1.0)
    h.set_scalar(P.S_CNM1, 0.0)
    m = P.CNAB2 if method == "CNAB2" else P.LS_IMEX_RK3
    step = O.cnab2 if method == "CNAB2" else O.rk3
    for _ in range(nsteps):
        step(S)
    h.step(nsteps, m, 0)
    om = h.download(P.F_OM)
    fn = h.download(P.F_FNM1)
    h.physical()
    ur = h.download(P.F_UR)
    e_om = rel_l2(om, S.om2)
    e_fn = rel_l2(fn, S.fnm1)
    e_u = rel_l2(ur[0:2], S.ur_full[0:2])
    print(f"  rel L2 om={e_om:.3e} fnm1={e_fn:.3e} u={e_u:.3e}", flush=True)
    assert e_om < 1e-10 and e_fn < 1e-10 and e_u < 1e-10


@pytest.mark.parametrize("N", [8, 16, 32, 64, 128, 2048])
def test_every_small_plan_and_variant_fp64(N):
    """Every compiled FFT plan/variant at the lengths the other tests use
    least (M = 12 .. 192, and 3072), FP64, 3 fixed-dt RK3 steps against the
    oracle: relative L2 <= 1e-10."""
    print(f"case: all variants N={N} fp64 3 RK3 steps", flush=True)
    S = O.create_state(N, 1e4, 3.0 if N < 32 else 10.0, 0.5, 4895, "KM3", dtype=np.float64)
    O.init_dt(S)
    dt = S.dt
    for _ in range(3):
        O.rk3(S)
    om0 = P.calcom_f32(S.uc0)
    for v in range(P.n_variants(N)):
        h = P.Handle(N, 8)
        h.set_variants(v, v)
        h.load_initial(S.uc0, om0)
        h.set_scalar(P.S_VISC, S.visc)
        h.set_scalar(P.S_DT, dt)
        h.step(3, P.LS_IMEX_RK3, 0)
        e = rel_l2(h.download(P.F_OM), S.om2)
        print(f"  variant {v}: rel L2 omega {e:.2e}", flush=True)
        assert e <= 1e-10
        h.close()


@pytest.mark.parametrize("N", [48, 96, 192, 384, 768, 1536, 3072])
def test_three_pow2_grids_fp64(N):
    """Grids N = 3*2^k (padded M = 9*2^(k-1), the second factor 3 inside the
    FFT plans): every compiled variant, FP64, 3 fixed-dt RK3 steps and then
    2 CNAB2 steps against the oracle, relative L2 <= 1e-10 on omega and u."""
    print(f"case: N={N} (3*2^k) fp64 3 RK3 + 2 CNAB2 steps", flush=True)
    # the host init equals the oracle's at these N (test_boundary); it saves the
    # oracle's pure-Python seeding loop at the larger grids
    uc0, _ = P.pao_generate(N, 10.0, 1e4, 4895, True)
    S = O.create_state(N, 1e4, 10.0, 0.5, 4895, "KM3", dtype=np.float64, uc0=uc0)
    O.init_dt(S)
    dt = S.dt
    for _ in range(3):
        O.rk3(S)
    for _ in range(2):
        O.cnab2(S)
    om0 = P.calcom_f32(S.uc0)
    for v in range(P.n_variants(N)):
        h = P.Handle(N, 8)
        h.set_variants(v, v)
        h.load_initial(S.uc0, om0)
        h.set_scalar(P.S_VISC, S.visc)
        h.set_scalar(P.S_DT, dt)
        h.step(3, P.LS_IMEX_RK3, 0)
        h.step(2, P.CNAB2, 0)
        e = rel_l2(h.download(P.F_OM), S.om2)
        h.physical()
        eu = rel_l2(h.download(P.F_UR)[0], S.ur_full[0])
        print(f"  variant {v}: rel L2 omega {e:.2e} u {eu:.2e}", flush=True)
        assert e <= 1e-10 and eu <= 1e-10
        h.close()


@pytest.mark.parametrize("N,prec", [(256, 8), (256, 4), (384, 4), (1024, 8), (4096, 4)])
def test_step_equals_stage_by_stage(N, prec):
    """RK3 stages run as one step (palx_step, or palx_stage_update with
    STAGE_IN_STEP) store the next stage's pre-update term omega (1 - K2 beta
    dt visc) + fnm1 beta dt in fnm1's storage at stages 0 and 1, and read it
    at stages 1 and 2 in place of omega and fnm1; plain stages keep omega and
    fnm1 after every stage.  Both form that term with the same fused
    multiply-adds, so two steps of each agree bit for bit on omega, fnm1 and
    the velocity, for every compiled variant (N=256, and N=384 whose 576-point
    plans carry a second factor 3) and the default ones (N=1024 FP64, N=4096
    FP32)."""
    print(f"case: palx_step / in-step stages vs plain stages N={N} prec={prec}", flush=True)
    uc0, st = P.pao_generate(N, 10.0, 1e4, 4895, True)
    om0 = P.calcom_f32(uc0)
    variants = [v for v in range(P.n_variants(N)) if P.variant_ok(N, prec, v)] if N <= 384 else [None]
    for v in variants:
        out = {}
        for how in ("plain", "step", "in_step"):
            h = P.Handle(N, prec)
            if v is not None:
                h.set_variants(v, v)
            h.load_initial(uc0, om0)
            h.set_scalar(P.S_VISC, st[4])
            h.set_scalar(P.S_DT, 2e-4 if N == 256 else 5e-5)
            for _ in range(2):
                if how == "step":
                    h.step(1, P.LS_IMEX_RK3, 0)
                else:
                    for s in range(3):
                        h.stage_products()
                        h.stage_update(P.LS_IMEX_RK3, s | (P.STAGE_IN_STEP if how == "in_step" else 0))
            om, fn = h.download(P.F_OM), h.download(P.F_FNM1)
            h.physical()
            out[how] = (om, fn, h.download(P.F_UR)[0:2])
            h.close()
        a_om, a_fn, a_u = out["plain"]
        assert np.abs(a_fn).max() > 0
        for how in ("step", "in_step"):
            b_om, b_fn, b_u = out[how]
            print(f"  variant {v} {how}: max diff om {np.abs(a_om - b_om).max():.1e} "
                  f"fn {np.abs(a_fn - b_fn).max():.1e} u {np.abs(a_u - b_u).max():.1e}", flush=True)
            assert np.array_equal(a_om, b_om) and np.array_equal(a_fn, b_fn) and np.array_equal(a_u, b_u)


def test_open_in_step_rk3_refuses_omega_access():
    """While an RK3 step run with STAGE_IN_STEP is open, omega / fnm1 hold
    intermediates: everything that reads or writes them fails with
    PALX_ESTATE (plain stages, palx_step, download / upload / field pointer
    of omega and fnm1, the omega getter, spectra, diagnostics), stages out of
    order fail, and the velocity stays readable.  The step closes at stage 2,
    and palx_load_initial closes an open one."""
    print("case: open in-step RK3 step refusals N=64", flush=True)
    N = 64
    uc0, st = P.pao_generate(N, 10.0, 1e4, 4895, True)
    om0 = P.calcom_f32(uc0)
    h = P.Handle(N, 4)
    h.load_initial(uc0, om0)
    h.set_scalar(P.S_VISC, st[4])
    h.set_scalar(P.S_DT, 1e-3)
    h.stage_products()
    with pytest.raises(P.PalxError):  # an in-step step starts at stage 0
        h.stage_update(P.LS_IMEX_RK3, 1 | P.STAGE_IN_STEP)
    with pytest.raises(ValueError):  # CNAB2 has no in-step form
        h.stage_update(P.CNAB2, 0 | P.STAGE_IN_STEP)
    h.stage_update(P.LS_IMEX_RK3, 0 | P.STAGE_IN_STEP)
    refused = [lambda: h.download(P.F_OM), lambda: h.download(P.F_FNM1),
               lambda: h.upload(P.F_OM, om0), lambda: h.step(1, P.LS_IMEX_RK3, 0),
               lambda: h.device_ptr(P.F_FNM1), lambda: h.spectrum(16), lambda: h.diag(),
               lambda: h.getter(P.VAR_OMEGA), lambda: h.recon_velocity()]
    for f in refused:
        with pytest.raises(P.PalxError) as ei:
            f()
        assert ei.value.code == P.PALX_ESTATE and "STAGE_IN_STEP" in str(ei.value)
    h.physical()
    assert np.isfinite(h.download(P.F_UR)[0:2]).all()
    h.stage_products()
    with pytest.raises(P.PalxError):  # a plain stage inside the open step
        h.stage_update(P.LS_IMEX_RK3, 1)
    with pytest.raises(P.PalxError):  # out of order
        h.stage_update(P.LS_IMEX_RK3, 2 | P.STAGE_IN_STEP)
    h.stage_update(P.LS_IMEX_RK3, 1 | P.STAGE_IN_STEP)
    h.stage_products()
    h.stage_update(P.LS_IMEX_RK3, 2 | P.STAGE_IN_STEP)
    assert np.isfinite(h.download(P.F_OM)).all()  # closed
    h.stage_products()
    h.stage_update(P.LS_IMEX_RK3, 0 | P.STAGE_IN_STEP)
    h.load_initial(uc0, om0)  # closes the open step
    assert np.array_equal(h.download(P.F_OM), om0)
    h.step(1, P.LS_IMEX_RK3, 0)
    h.close()
    # an ensemble refuses to step or tune while one of its members has an open step
    from palinstrophy_amd.ensemble import DnsEnsemble, Member
    ens = DnsEnsemble([Member(64, 4895, 10.0, "KM3"), Member(64, 4896, 10.0, "PAO")], precision=4, tune=False)
    m0 = ens.handles[0]
    m0.stage_products()
    m0.stage_update(P.LS_IMEX_RK3, 0 | P.STAGE_IN_STEP)
    for f in (lambda: ens.step(1, P.LS_IMEX_RK3, 0), lambda: ens.ens.tune(1)):
        with pytest.raises(P.PalxError) as ei:
            f()
        assert ei.value.code == P.PALX_ESTATE
    for s in (1, 2):
        m0.stage_products()
        m0.stage_update(P.LS_IMEX_RK3, s | P.STAGE_IN_STEP)
    ens.step(1, P.LS_IMEX_RK3, 0)
    ens.close()
