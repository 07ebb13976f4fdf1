"""The reference GUI's own expressions on this state (no PySide6 needed).

Each block below applies, to a DnsState of this build, the expressions the
reference callers apply to theirs (turbo_main.py line numbers cited), with
the state's ``xp`` in the role of CuPy, and checks the result against the
library's own diagnostics / downloads."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(N=64, precision="fp64", seed=7):
    from palinstrophy_amd import turbo_simulator as dns_all
    S = dns_all.create_dns_state(N=N, Re=1e4, K0=10.0, CFL=0.5, backend="gpu", seed=seed,
                                 start_spectrum="KM3", precision=precision)
    dns_all.dns_step2a(S)
    S.dt = S.cflnum / (float(dns_all.compute_cflm(S)) * math.pi)
    for _ in range(4):
        dns_all.dns_step_ls_imex_rk3(S)
    return dns_all, S


def test_field_fetch_get_and_asarray():
    """turbo_main.py:900-921: xp.asarray(S.ur_full[k], dtype=xp.float32), then
    field.get(); turbo_main.py:1280-1281 / 2061-2065: np.asarray on device data."""
    from palinstrophy_amd import _palx as P
    print("case: callers field fetch N=64 fp64", flush=True)
    dns_all, S = _state()
    xp = S.xp
    field = xp.asarray(S.ur_full[0], dtype=xp.float32)
    host = (field.get() if S.backend == "gpu" else field).astype(np.float32)
    S.handle.physical()
    ur = S.handle.download(P.F_UR)
    assert host.dtype == np.float32 and host.shape == ur[0].shape
    assert np.array_equal(host, ur[0].astype(np.float32))
    assert np.array_equal(np.asarray(S.om2), S.handle.download(P.F_OM))
    dns_all.dns_om2_phys(S)
    assert np.asarray(xp.asarray(S.ur_full[2], dtype=xp.float32)).shape == ur[0].shape


def test_energy_joules_and_pal_over_ens_kmax2_expressions():
    """turbo_main.py:1947-1959 and 2015-2073 on the state, against palx_diag."""
    print("case: callers metrics N=64 fp64", flush=True)
    dns_all, S = _state()
    xp = S.xp
    # energy_joules
    u = S.ur_full[0]
    v = S.ur_full[1]
    velocity_sq_sum = xp.sum(u * u, dtype=xp.float64) + xp.sum(v * v, dtype=xp.float64)
    value = 0.5 * (2.0 * math.pi) ** 2 * velocity_sq_sum / u.size
    if hasattr(value, "get"):
        value = value.get()
    e_j = float(value)
    # pal_over_ens_kmax2
    band = S.om2
    K2 = S.step3_K2
    NX_half = int(band.shape[1])
    w = getattr(S, "_pal_rfft_w", None)
    if w is None or w.shape[0] != NX_half:
        w_np = np.empty(NX_half, dtype=np.float32)
        w_np[:] = 2.0
        w_np[0] = 1.0
        if NX_half > 1:
            w_np[-1] = 1.0
        w = xp.asarray(w_np)
        S._pal_rfft_w = w
    P_ = band.real * band.real + band.imag * band.imag
    WP = w * P_
    buf = xp.stack([WP.sum(), P_[0, 0], K2.max(), (w * K2 * P_).sum()])
    vals = buf.get() if hasattr(buf, "get") else np.asarray(buf)
    total = float(vals[0] - vals[1])
    ratio = float(vals[3]) / (total * float(vals[2]))
    d = S.diagnostics()
    assert e_j == pytest.approx(d["energy_joules"], rel=1e-12)
    assert total == pytest.approx(d["enstrophy"], rel=1e-12)
    assert float(vals[3]) == pytest.approx(d["palinstrophy"], rel=1e-12)
    assert float(vals[2]) == d["kmax2"]
    assert ratio == pytest.approx(d["pal_over_ens_kmax2"], rel=1e-12)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_restart_dump_and_load_expressions(precision):
    """turbo_main.py:1329-1336 (dump: S.uc, S.om2, S.fnm1 to host) and
    1412-1450 (load: S.uc = xp.asarray(...), S.uc_full scatter with
    xp.transpose(S.uc, (2, 0, 1)), dns_step2a).  The loaded state continues
    like the original: its velocity comes from uc, which holds the velocity
    reconstructed from the current omega (complex64, as the file stores it)."""
    from palinstrophy_amd import restart
    print(f"case: callers restart N=64 {precision}", flush=True)
    dns_all, S = _state(precision=precision)
    host = lambda arr: np.asarray(arr.get() if hasattr(arr, "get") else np.asarray(arr))  # noqa: E731
    uc_data = host(S.uc).astype(np.complex64)
    om2_data = host(S.om2).astype(np.complex64)
    fnm1_data = host(S.fnm1).astype(np.complex64)
    N, NK = S.Nbase, S.NK
    assert uc_data.shape == (N, NK, 3)
    ref_uc = restart.velocity_from_omega(host(S.om2))
    assert np.abs(uc_data - ref_uc).max() <= 2e-6 * np.abs(ref_uc).max()
    assert not np.any(uc_data[:, N // 2:, :]) and not np.any(uc_data[:, :, 2])
    # a second state with the saved case loaded the way the GUI does it
    S2 = dns_all.create_dns_state(N=N, Re=S.Re, K0=S.K0, CFL=S.cflnum, backend="gpu", seed=S.seed_init,
                                  skip_pao=True, start_spectrum="KM3", precision=precision)
    xp = S2.xp
    S2.uc = xp.asarray(uc_data)
    S2.om2 = xp.asarray(om2_data)
    S2.fnm1 = xp.asarray(fnm1_data)
    S2.visc, S2.cflnum, S2.t, S2.dt, S2.cn, S2.cnm1 = S.visc, S.cflnum, S.t, S.dt, S.cn, S.cnm1
    S2.uc_full[...] = 0
    S2.uc_full[:, :S2.NZ, :NK] = xp.transpose(S2.uc, (2, 0, 1))
    dns_all.dns_step2a(S2)
    # same physical velocity as the original state (to complex64 rounding of uc)
    u1, u2 = host(S.ur_full[0:2]), host(S2.ur_full[0:2])
    assert np.abs(u2 - u1).max() <= 3e-6 * np.abs(u1).max()
    # and the two runs continue together
    S.om2 = xp.asarray(om2_data)  # the same complex64 omega in both
    for _ in range(5):
        dns_all.dns_step_ls_imex_rk3(S)
        dns_all.dns_step_ls_imex_rk3(S2)
    a, b = host(S.om2), host(S2.om2)
    assert np.linalg.norm(b - a) <= 1e-5 * np.linalg.norm(a)
    # reading uc_full hands out the current velocity; dns_step2a then reloads
    # that same velocity
    before = host(S2.ur_full[0:2])
    _ = S2.uc_full
    dns_all.dns_step2a(S2)
    after = host(S2.ur_full[0:2])
    assert np.abs(after - before).max() <= 3e-6 * np.abs(before).max()
