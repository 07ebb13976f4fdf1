"""GPU test of the restart round trip: a run interrupted by
dump_restart_netcdf / load_restart_netcdf continues exactly like the
uninterrupted run (bit-identical in FP32, whose state the file stores
exactly; within 1e-14 in FP64 through the float64 extras)."""
import numpy as np
import pytest

from palinstrophy_amd import _palx as P
from palinstrophy_amd import restart as R
from palinstrophy_amd import turbo_simulator as ts

pytestmark = pytest.mark.gpu


def start(precision):
    S = ts.create_dns_state(N=64, Re=1e4, K0=10.0, CFL=0.5, backend="gpu", seed=4895,
                            start_spectrum="KM3", precision=precision)
    S.dt = 0.5 / (ts.compute_cflm(S) * np.pi)
    return S


def advance(S, nsteps, first_it):
    S.handle.set_iteration(first_it)
    S.handle.step(nsteps, P.LS_IMEX_RK3, 5)
    S.it = first_it + nsteps
    S.t = float(S.handle.scalars()[P.S_T])


@pytest.mark.parametrize("precision,tol", [("fp32", 0.0), ("fp64", 1e-14)])
def test_restart_continues_the_run(tmp_path, precision, tol):
    print(f"case: restart round trip N=64 {precision}", flush=True)
    A = start(precision)
    advance(A, 20, 0)
    ref = A.handle.download(P.F_OM)
    ref_sc = A.handle.scalars()

    B = start(precision)
    advance(B, 10, 0)
    R.dump_restart_netcdf(B, str(tmp_path))
    C = R.load_restart_netcdf(str(tmp_path), precision=precision)
    assert C.it == 10 and C.Nbase == 64
    advance(C, 10, 10)
    om = C.handle.download(P.F_OM)
    sc = C.handle.scalars()
    e = float(np.linalg.norm(om - ref) / np.linalg.norm(ref))
    print(f"  rel L2 omega {e:.2e}; dt {sc[P.S_DT]:.17e} vs {ref_sc[P.S_DT]:.17e}", flush=True)
    assert e <= tol
    assert abs(sc[P.S_DT] - ref_sc[P.S_DT]) <= tol * ref_sc[P.S_DT]
    assert abs(sc[P.S_T] - ref_sc[P.S_T]) <= 1e-14 * ref_sc[P.S_T]


def test_restart_file_against_oracle(tmp_path):
    """FP64: the file written at step 10 restarts both a new handle and the
    FP64 oracle (omega, fnm1 and the scalars read back from the file); the two
    continue 10 adaptive steps together to 1e-10.  The file's uc is the
    oracle's velocity reconstruction of the stored omega (turbo_simulator.py:
    1555-1585) to complex64 rounding, and the restarted handle's physical
    velocity equals the oracle's padded inverse of it."""
    from oracle import dns_oracle as O
    print("case: restart vs oracle N=64 fp64", flush=True)
    B = start("fp64")
    advance(B, 10, 0)
    R.dump_restart_netcdf(B, str(tmp_path))
    attrs, arrays = R.read_restart_file(str(tmp_path / R.RESTART_FILE))
    assert arrays["om2"].dtype == np.complex128  # the float64 extras of an FP64 state
    N = int(attrs["Nbase"])
    S = O.create_state(N, attrs["Re"], attrs["K0"], attrs["cflnum"], int(attrs["seed_init"]),
                       attrs["start_spectrum"], dtype=np.float64)
    S.arrays["om2"] = arrays["om2"].copy()
    S.arrays["fnm1"] = arrays["fnm1"].copy()
    S.visc, S.dt, S.cn, S.cnm1, S.t = attrs["visc"], attrs["dt"], attrs["cn"], attrs["cnm1"], attrs["t"]
    S.uc_full[...] = 0
    O.reconstruct(S)
    H = N // 2
    uc = arrays["uc"]
    rec = S.uc_full[0:2, :N, :H]
    assert np.abs(uc[:, :H, 0] - rec[0]).max() <= 1e-6 * np.abs(rec[0]).max()
    assert np.abs(uc[:, :H, 1] - rec[1]).max() <= 1e-6 * np.abs(rec[1]).max()
    O.step2a(S)
    C = R.load_restart_netcdf(str(tmp_path), precision="fp64")
    C.handle.physical()
    ur = C.handle.download(P.F_UR)
    assert np.abs(ur[0:2] - S.ur_full[0:2]).max() <= 1e-12 * np.abs(S.ur_full[0:2]).max()
    advance(C, 10, 10)
    for it in range(11, 21):  # palx_step's cadence from it = 10 with nextdt_every = 5
        O.rk3(S)
        if it % 5 == 0:
            O.next_dt(S)
    om = C.handle.download(P.F_OM)
    e = float(np.linalg.norm(om - S.om2) / np.linalg.norm(S.om2))
    sc = C.handle.scalars()
    print(f"  rel L2 omega vs oracle {e:.2e}; dt {sc[P.S_DT]:.16e} vs {S.dt:.16e}", flush=True)
    assert e <= 1e-10
    assert abs(sc[P.S_DT] - S.dt) <= 1e-12 * S.dt
