"""The CHECK time-stepper comparison (turbo_simulator.py:2868-2903) against
the same two runs of the FP64 oracle."""
import io
import contextlib
import re

import numpy as np
import pytest

from oracle import dns_oracle as O

pytestmark = pytest.mark.gpu


def _rel(ref, cand):
    ref = np.asarray(ref).astype(np.complex128)
    cand = np.asarray(cand).astype(np.complex128)
    d = cand - ref
    return (float(np.linalg.norm(d.ravel()) / max(np.linalg.norm(ref.ravel()), 1e-30)),
            float(np.max(np.abs(d)) / max(float(np.max(np.abs(ref))), 1e-30)))


def test_check_method_matches_oracle(monkeypatch):
    from palinstrophy_amd import turbo_simulator as ts
    print("case: CHECK N=64 KM3 fp64", flush=True)
    N, Re, K0, STEPS, CFL, UPDATE, seed = 64, 1e4, 5.0, 40, 0.5, 10, 4895
    monkeypatch.setenv("SCIPYTURBO_SEED", str(seed))
    out = io.StringIO()
    with contextlib.redirect_stdout(out):
        got = ts.compare_time_steppers(N, Re, K0, STEPS, CFL, "gpu", "KM3", UPDATE, "fp64")
        assert ts.run_dns(N, Re, K0, STEPS, CFL, "gpu", "KM3", UPDATE, "CHECK", "fp64") is None
    a, _ = O.run(N, Re, K0, STEPS, CFL, "KM3", UPDATE, "CNAB2", seed=seed, dtype=np.float64)
    b, _ = O.run(N, Re, K0, STEPS, CFL, "KM3", UPDATE, "LS_IMEX_RK3", seed=seed, dtype=np.float64)
    om_l2, om_linf = _rel(a.om2, b.om2)
    u_l2, u_linf = _rel(a.ur_full[0:2], b.ur_full[0:2])
    e0, _, tau0 = O.flow_eddy_metrics(a)
    e1, _, tau1 = O.flow_eddy_metrics(b)
    want = {"om_l2": om_l2, "om_linf": om_linf, "u_l2": u_l2, "u_linf": u_linf,
            "e_cnab2": e0, "e_rk3": e1, "tau_cnab2": tau0, "tau_rk3": tau1}
    assert om_l2 > 1e-8  # the two steppers do differ
    for k, v in want.items():
        assert got[k] == pytest.approx(v, rel=1e-9), k
    # both CHECK RESULT blocks print the same figures, to their printed digits
    text = out.getvalue()
    blocks = re.findall(r"OM2 relative L2=(\S+) Linf=(\S+)\n U   relative L2=(\S+) Linf=(\S+)", text)
    assert len(blocks) == 2
    for blk in blocks:
        for s, k in zip(blk, ("om_l2", "om_linf", "u_l2", "u_linf")):
            assert float(s) == pytest.approx(want[k], rel=2e-6)
    assert text.count("--- CHECK RESULT ---") == 2


def test_config1_cli_command_runs_on_the_gpu(monkeypatch, capsys):
    """`sim 256 10000 10 1001 0.75 cpu 100 KM3` (the reference's config-1 command
    line, SCIPYTURBO_SEED=4895): the `cpu` backend runs on the GPU and says so;
    the ITERATION records at 1, 100, 200 match the reference's own
    (tests/golden), and at it = 300, where the reference prints NaN, the
    facade's NEXTDT sync raises FloatingPointError (FP64, 300 steps)."""
    import json
    import math
    import os
    from palinstrophy_amd import turbo_simulator as ts
    print("case: sim 256 10000 10 300 0.75 cpu 100 KM3 fp64", flush=True)
    monkeypatch.setenv("SCIPYTURBO_SEED", "4895")
    monkeypatch.setenv("PALINSTROPHY_PRECISION", "fp64")
    with np.errstate(all="ignore"), pytest.raises(FloatingPointError):
        ts.main(["256", "10000", "10", "300", "0.75", "cpu", "100", "KM3"])
    out = capsys.readouterr().out
    assert "backend 'cpu' is not built: running on the GPU" in out
    recs = {int(m.group(1)): [float(m.group(k)) for k in range(2, 7)]
            for m in re.finditer(r"ITERATION\s+(\d+)\s+T=\s*(\S+)\s+DT=\s*(\S+)\s+CN=\s*(\S+)\s+CFLM=\s*(\S+)\s+TAU_L=\s*(\S+)", out)}
    with open(os.path.join(os.path.dirname(__file__), "golden", "config1_N256_km3_cnab2.json")) as f:
        g = {int(r[0]): r[1:] for r in json.load(f)["records"]}
    assert sorted(recs) == [1, 100, 200]
    for it in (1, 100, 200):
        for got, want in zip(recs[it], g[it]):
            assert abs(got - want) <= 2e-5 * abs(want) + 1e-9, (it, got, want)
    assert not math.isfinite(g[300][3])
