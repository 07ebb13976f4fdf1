"""Write golden vectors for the functions no reference log pins.

Runs the reference module ``palinstrophy/turbo_simulator.py`` (CPU backend,
pure-Python init since numba is absent) from /root/reference, in the build
container only, and stores small .npz fixtures next to this script:

  km3_rk3_N{N}_s{seed}.npz : KM3 initial spectra (uc_full[0:2,:N,:N/2]) and
      omega_0; then 10 fixed-dt LS-IMEX-RK3 steps: omega, fnm1, ur_full[0:2],
      energy-spectrum bins, dns_om2_phys / dns_stream_func / dns_kinetic
      planes, flow_eddy_metrics (physical planes only for N=32).
  km3_cnab2_N32_s7.npz : KM3 + 10 fixed-dt CNAB2 steps (omega, ur_full[0:2]).

The PAO initial field and the CNAB2 + NEXTDT trajectory are pinned by the
reference's recorded logs instead (log_sim_512_1001.json, extract_logs.py).

Usage: python tests/golden/make_golden.py /root/reference
"""
import contextlib
import io
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref_root):
    sys.path.insert(0, ref_root)
    with contextlib.redirect_stdout(io.StringIO()):
        from palinstrophy import turbo_simulator as ts  # noqa: E402

    def run(N, seed, method, steps, Re=1e4, K0=5.0, cfl=0.5):
        with contextlib.redirect_stdout(io.StringIO()):
            S = ts.create_dns_state(N=N, Re=Re, K0=K0, CFL=cfl, backend="cpu", seed=seed,
                                    start_spectrum="KM3", populate_compact_ur=False)
            H = N // 2
            uc0 = np.array(S.uc_full[0:2, :N, :H])
            om0 = np.array(S.om2)
            ts.dns_step2a(S)
            cflm0 = float(ts.compute_cflm(S))
            S.dt = S.cflnum / (cflm0 * np.pi)
            S.cn, S.cnm1 = 1.0, 0.0
            for _ in range(steps):
                if method == "RK3":
                    ts.dns_step_ls_imex_rk3(S)
                else:
                    ts.dns_step2b(S)
                    ts.dns_step3(S)
                    ts.dns_step2a(S)
            out = dict(N=N, seed=seed, Re=Re, K0=K0, cfl=cfl, steps=steps, dt=S.dt, cflm0=cflm0,
                       visc=float(S.visc), uc0=uc0, om0=om0, om=np.array(S.om2),
                       fnm1=np.array(S.fnm1), ur=np.array(S.ur_full[0:2]))
            if method == "RK3":
                U, L, tau = S.flow_eddy_metrics()
                out.update(U=U, L=L, tau=tau)
                sp = ts.compute_energy_spectrum_uv_bins(S)
                out.update(spec_e=sp.energy, spec_c=sp.count)
                ts.dns_om2_phys(S)
                out["omega_phys"] = np.array(S.ur_full[2])
                ts.dns_stream_func(S)
                out["psi_phys"] = np.array(S.ur_full[2])
                ts.dns_kinetic(S)
                out["ke_phys"] = np.array(S.ur_full[2])
        return out

    for N, seed in ((32, 1), (32, 4895), (64, 4895)):
        d = run(N, seed, "RK3", 10)
        if N > 32:  # keep the fixture small: spectral fields and scalars only
            for k in ("ur", "omega_phys", "psi_phys", "ke_phys"):
                d.pop(k)
        np.savez_compressed(os.path.join(HERE, f"km3_rk3_N{N}_s{seed}.npz"), **d)
        print("wrote", N, seed)
    d = run(32, 7, "CNAB2", 10)
    np.savez_compressed(os.path.join(HERE, "km3_cnab2_N32_s7.npz"), **d)
    print("wrote cnab2")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
