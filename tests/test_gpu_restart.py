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
