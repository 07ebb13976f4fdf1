"""Golden records of the `sim 256 10000 10 1001 0.75 cpu 100 KM3` run (CNAB2).

Runs the reference's own ``run_dns`` loop (palinstrophy/turbo_simulator.py,
CPU backend) from the reference tree, in the build container only, with
SCIPYTURBO_SEED=4895, and stores next to this script:

  config1_N256_km3_cnab2.json : the ITERATION records the loop prints (it, T,
      DT, CN, CFLM, TAU_L; "nan" where the field has blown up) and the first
      record with a non-finite CFLM;
  config1_N256_step200.npz : omega after 200 steps on the modes |kz| < 16,
      kx < 16 (complex64), its full L2 norm and the enstrophy sum.

Usage: python -B tests/golden/make_config1.py /root/reference
"""
import contextlib
import io
import json
import math
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REC = re.compile(r"ITERATION\s+(\d+)\s+T=\s*(\S+)\s+DT=\s*(\S+)\s+CN=\s*(\S+)\s+CFLM=\s*(\S+)\s+TAU_L=\s*(\S+)")


def main(ref_root):
    sys.dont_write_bytecode = True
    sys.path.insert(0, ref_root)
    os.environ["SCIPYTURBO_SEED"] = "4895"
    with contextlib.redirect_stdout(io.StringIO()):
        from palinstrophy import turbo_simulator as ts  # noqa: E402
    args = dict(N=256, Re=10000, K0=10.0, CFL=0.75, backend="cpu", start_spectrum="KM3",
                UPDATE=100, method="CNAB2")
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf), np.errstate(all="ignore"):
        ts.run_dns(STEPS=1001, **args)
    recs = [[int(m.group(1))] + [float(m.group(k)) for k in range(2, 7)] for m in REC.finditer(buf.getvalue())]
    first_bad = next((r[0] for r in recs if not math.isfinite(r[4])), None)
    with open(os.path.join(HERE, "config1_N256_km3_cnab2.json"), "w") as f:
        json.dump({"command": "sim 256 10000 10 1001 0.75 cpu 100 KM3 (SCIPYTURBO_SEED=4895)",
                   "fields": ["it", "T", "DT", "CN", "CFLM", "TAU_L"], "records": recs,
                   "first_nonfinite_it": first_bad}, f, indent=1)
    with contextlib.redirect_stdout(io.StringIO()), np.errstate(all="ignore"):
        S = ts.run_dns(STEPS=200, **args)
    om = np.asarray(S.om2).astype(np.complex128)
    N = om.shape[0]
    rows = np.r_[0:16, N - 15:N]
    w = np.full(om.shape[1], 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    np.savez_compressed(os.path.join(HERE, "config1_N256_step200.npz"),
                        rows=rows, om_low=om[rows, :16].astype(np.complex64),
                        norm=np.linalg.norm(om), enstrophy=float(np.sum(w * np.abs(om) ** 2) - abs(om[0, 0]) ** 2),
                        t=float(S.t), dt=float(S.dt))
    print("records", len(recs), "first non-finite", first_bad)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
