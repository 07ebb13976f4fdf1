"""GPU tests of the kernel variants, the autotuner and the ensemble form.

* every compiled variant of the row / column step passes meets the FP64 bar
  (relative L2 <= 1e-10 against the oracle after fixed-dt steps);
* an ensemble member is bit-identical to the same run on its own handle
  (same kernels, same variants), with the graph, eager-forked and serial
  launch orders all giving the same bits;
* one member of an adaptive-dt ensemble follows the oracle's trajectory
  (NEXTDT cadence included) to the FP64 bar;
* the autotuner picks valid variants and reuses its cache.
"""
import math

import numpy as np
import pytest

from oracle import dns_oracle as O
from palinstrophy_amd import _palx as P
from palinstrophy_amd import autotune
from palinstrophy_amd.ensemble import DnsEnsemble, Member

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a).astype(np.complex128).ravel()
    b = np.asarray(b).astype(np.complex128).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("N,method", [(256, "RK3"), (512, "RK3"), (1024, "RK3"), (256, "CNAB2")])
def test_every_variant_meets_fp64_bar(N, method):
    print(f"case: variants N={N} {method} fp64 4 fixed-dt steps", flush=True)
    S = O.create_state(N, 1e4, 10.0, 0.5, 4895, "KM3", dtype=np.float64)
    O.init_dt(S)
    dt = S.dt
    step = O.rk3 if method == "RK3" else O.cnab2
    for _ in range(4):
        step(S)
    nv = P.n_variants(N)
    assert nv == P.MAX_VARIANTS
    mid = P.LS_IMEX_RK3 if method == "RK3" else P.CNAB2
    om0 = P.calcom_f32(S.uc0)
    for vr in range(nv):
        for vc in range(nv):
            if vr != vc and vr != 0:  # every row and every column variant
                continue
            h = P.Handle(N, 8)
            h.set_variants(vr, vc)
            assert h.variants() == (vr, vc)
            h.load_initial(S.uc0, om0)
            h.set_scalar(P.S_VISC, S.visc)
            h.set_scalar(P.S_DT, dt)
            h.set_scalar(P.S_CN, 1.0)
            h.set_scalar(P.S_CNM1, 0.0)
            h.step(4, mid, 0)
            e = rel_l2(h.download(P.F_OM), S.om2)
            print(f"  row v{vr} col v{vc}: rel L2 omega {e:.2e}", flush=True)
            assert e <= 1e-10
            h.close()


def test_variant_out_of_range_is_rejected():
    h = P.Handle(64, 8)
    with pytest.raises(ValueError):
        h.set_variants(P.MAX_VARIANTS, 0)
    with pytest.raises(ValueError):
        h.set_variants(0, 3)  # N=64 has the three layouts of the base plan only
    h.close()


@pytest.mark.parametrize("N,prec", [(16384, 4), (8192, 8), (12288, 8)])
def test_largest_grids_start_on_and_keep_fitting_variants(N, prec):
    """Where a complex transform does not fit a block's LDS (N=16384 FP32,
    N=8192 and 12288 FP64), a handle starts on split-exchange variants and
    refuses the others (no launch ever fails on an oversized block)."""
    print(f"case: fitting variants N={N} prec={prec}", flush=True)
    h = P.Handle(N, prec)
    vr, vc = h.variants()
    assert P.variant_ok(N, prec, vr) and P.variant_ok(N, prec, vc) and vr >= 6 and vc >= 6
    with pytest.raises(ValueError):
        h.set_variants(0, vc)
    with pytest.raises(ValueError):
        h.set_variants(vr, 1)
    assert h.variants() == (vr, vc)
    h.close()


MEMBERS = [Member(32, 1, 10.0, "KM3", 1e4), Member(64, 2, 5.0, "PAO", 5e3), Member(32, 3, 15.0, "PAO", 1e4),
           Member(128, 4, 10.0, "KM3", 1e4), Member(64, 5, 15.0, "KM3", 2e4), Member(48, 6, 5.0, "KM3", 1e4),
           Member(96, 7, 10.0, "PAO", 1e4), Member(48, 8, 10.0, "PAO", 5e3)]


def standalone(m, precision, vr, vc, nsteps, method, update):
    h = P.Handle(m.N, precision)
    h.set_variants(vr, vc)
    uc0, st = P.pao_generate(m.N, m.K0, m.reynolds(), m.seed, m.spectrum == "KM3")
    h.load_initial(uc0, P.calcom_f32(uc0))
    h.set_scalar(P.S_VISC, st[4])
    h.set_scalar(P.S_CFLNUM, 0.5)
    h.set_scalar(P.S_CN, 1.0)
    h.set_scalar(P.S_CNM1, 0.0)
    h.set_scalar(P.S_DT, 0.5 / (h.cflm() * math.pi))
    h.step(nsteps, method, update)
    return h


@pytest.mark.parametrize("method", [P.LS_IMEX_RK3, P.CNAB2])
def test_ensemble_members_bit_identical_to_standalone(method):
    print(f"case: ensemble of 8 (N 32..128, with 3*2^k grids) fp64 method {method}, 12 steps NEXTDT every 5, "
          "21 steps NEXTDT every 10", flush=True)
    runs = {}
    for mode in ("graph", "fork", "serial"):
        ens = DnsEnsemble(MEMBERS, precision=8, CFL=0.5, tune=False)
        # graph mode runs 8 steps without NEXTDT in one launch: it 13..33 with
        # NEXTDT at 20 and 30 gives a chunk (21..28) and single steps around it
        for n, upd in ((12, 5), (21, 10)):
            if mode == "graph":
                ens.step(n, method, upd, graph=True)
            else:
                ens.ens.step(n, method, upd, graph=False, serial=(mode == "serial"))
        ens.sync()
        runs[mode] = [(h.download(P.F_OM), h.download(P.F_FNM1), h.scalars()) for h in ens.handles]
        ens.close()
    for i, m in enumerate(MEMBERS):
        h = standalone(m, 8, 0, 1, 12, method, 5)
        h.step(21, method, 10)
        om, fn, sc = h.download(P.F_OM), h.download(P.F_FNM1), h.scalars()
        for mode, res in runs.items():
            assert np.array_equal(res[i][0], om), (mode, i)
            assert np.array_equal(res[i][1], fn), (mode, i)
            for k in (P.S_DT, P.S_T, P.S_CN, P.S_CNM1, P.S_CFLM, P.S_IT):
                assert res[i][2][k] == sc[k], (mode, i, k)
        h.close()
    print("  graph == fork == serial == standalone (bit-exact)", flush=True)


def test_ensemble_member_follows_oracle_adaptive_fp64():
    print("case: ensemble member N=64 PAO vs oracle, 20 RK3 steps NEXTDT every 5", flush=True)
    ens = DnsEnsemble(MEMBERS, precision=8, CFL=0.5, tune=False)
    ens.step(20, P.LS_IMEX_RK3, 5)
    ens.sync()
    m = MEMBERS[1]
    S = O.create_state(m.N, m.reynolds(), m.K0, 0.5, m.seed, m.spectrum, dtype=np.float64)
    O.init_dt(S)
    for it in range(1, 21):
        O.rk3(S)
        if it == 1 or it % 5 == 0:
            O.next_dt(S)
    h = ens.handles[1]
    e = rel_l2(h.download(P.F_OM), S.om2)
    sc = h.scalars()
    print(f"  rel L2 omega {e:.2e}; dt {sc[P.S_DT]:.16e} vs {S.dt:.16e}", flush=True)
    assert e <= 1e-10
    assert abs(sc[P.S_DT] - S.dt) <= 1e-12 * S.dt
    ens.close()


def test_autotuner_valid_and_cached(tmp_path, monkeypatch):
    print("case: autotune ensemble 4xN256 + 2xN512 fp32", flush=True)
    monkeypatch.setenv("PALX_TUNE_CACHE", str(tmp_path / "tune.json"))
    mem = [Member(256, s, 10.0, "KM3") for s in range(1, 5)] + [Member(512, s, 10.0, "PAO") for s in (5, 6)]
    e1 = DnsEnsemble(mem, precision=4, tune=True, tune_reps=3)
    for N, r in e1.tuning.items():
        assert not r["cached"]
        nv = P.n_variants(N)
        assert 0 <= r["vrow"] < nv and 0 <= r["vcol"] < nv
        assert all(x > 0 for x in r["ms_row"][:nv]) and all(x > 0 for x in r["ms_col"][:nv])
        assert r["ms_row"][r["vrow"]] == min(r["ms_row"][:nv])
        print(f"  N={N}: row {r['ms_row']} -> v{r['vrow']}; col {r['ms_col']} -> v{r['vcol']}", flush=True)
    cls = {N: (vr, vc) for N, _, vr, vc in e1.ens.classes()}
    for N, r in e1.tuning.items():
        assert cls[N] == (r["vrow"], r["vcol"])
    e1.step(3, P.LS_IMEX_RK3, 100)
    e1.sync()
    for h in e1.handles:
        assert np.isfinite(h.scalars()[P.S_DT])
    e1.close()
    e2 = DnsEnsemble(mem, precision=4, tune=True, tune_reps=3)
    assert all(r["cached"] for r in e2.tuning.values())
    e2.close()
    h = P.Handle(512, 4)
    r = autotune.tune_handle(h, reps=3)
    assert h.variants() == (r["vrow"], r["vcol"])
    h.close()


@pytest.mark.parametrize("method", [P.LS_IMEX_RK3, P.CNAB2])
def test_single_handle_graph_stepping_bit_identical(method):
    print(f"case: palx_set_graph N=64 fp64 method {method}", flush=True)
    m = MEMBERS[1]
    res = []
    for graph in (False, True):
        h = P.Handle(m.N, 8)
        h.set_graph(graph)
        uc0, st = P.pao_generate(m.N, m.K0, m.reynolds(), m.seed, m.spectrum == "KM3")
        h.load_initial(uc0, P.calcom_f32(uc0))
        h.set_scalar(P.S_VISC, st[4])
        h.set_scalar(P.S_CFLNUM, 0.5)
        h.set_scalar(P.S_DT, 0.5 / (h.cflm() * math.pi))
        h.step(7, method, 5)
        h.set_variants(0, 0)  # a variant change drops the captured graphs
        h.step(5, method, 5)
        # graph mode runs 8 steps without NEXTDT in one launch: it 13..33 with
        # NEXTDT at 20 and 30 gives a chunk (21..28) and single steps around it
        h.step(21, method, 10)
        res.append((h.download(P.F_OM), h.download(P.F_FNM1), h.scalars()))
        h.close()
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    assert np.array_equal(res[0][2], res[1][2])


@pytest.mark.timeout(400)
def test_config4_mixed_ensemble_fp32(tmp_path, monkeypatch):
    """The batched FP32 ensemble workload: 64 x N256 + 16 x N512 + 4 x N1024
    (seeds 1..84, K0 from {5, 10, 15}, KM3 or PAO), tuned variants, 20 RK3
    steps in one graph launch per step, NEXTDT at it = 1 and every 10.  Every
    member is bit-identical to the same run on its own handle with the same
    variants; one member per size class is as close to the FP64 oracle as the
    FP32 oracle is (within 3x)."""
    from palinstrophy_amd.ensemble import mixed_ensemble
    print("case: config-4 ensemble 64xN256+16xN512+4xN1024 fp32, 20 RK3 steps, graph", flush=True)
    monkeypatch.setenv("PALX_TUNE_CACHE", str(tmp_path / "tune.json"))
    members = mixed_ensemble()
    assert len(members) == 84 and {m.N for m in members} == {256, 512, 1024}
    ens = DnsEnsemble(members, precision=4, CFL=0.5, tune=True, tune_reps=3)
    variants = {N: (vr, vc) for N, _, vr, vc in ens.ens.classes()}
    ens.step(20, P.LS_IMEX_RK3, 10, graph=True)
    ens.sync()
    res = [(h.download(P.F_OM), h.download(P.F_FNM1), h.scalars()) for h in ens.handles]
    ens.close()
    for i, m in enumerate(members):
        vr, vc = variants[m.N]
        h = standalone(m, 4, vr, vc, 20, P.LS_IMEX_RK3, 10)
        assert np.array_equal(res[i][0], h.download(P.F_OM)), i
        assert np.array_equal(res[i][1], h.download(P.F_FNM1)), i
        sc = h.scalars()
        for k in (P.S_DT, P.S_T, P.S_CN, P.S_CNM1):
            assert res[i][2][k] == sc[k], (i, k)
        h.close()
    print("  84 members bit-identical to their standalone runs", flush=True)
    first = {}
    for i, m in enumerate(members):
        first.setdefault(m.N, i)
    for N, i in sorted(first.items()):
        m = members[i]
        oracle = {}
        for dt in (np.float32, np.float64):
            S = O.create_state(m.N, m.reynolds(), m.K0, 0.5, m.seed, m.spectrum, dtype=dt)
            O.init_dt(S)
            for it in range(1, 21):
                O.rk3(S)
                if it == 1 or it % 10 == 0:
                    O.next_dt(S)
            oracle[dt] = S.om2
        e_gpu = rel_l2(res[i][0], oracle[np.float64])
        e_o32 = rel_l2(oracle[np.float32], oracle[np.float64])
        print(f"  N={N} member {i} ({m.spectrum}, K0={m.K0}): FP32 GPU vs FP64 oracle {e_gpu:.2e}, "
              f"FP32 oracle vs FP64 oracle {e_o32:.2e}", flush=True)
        assert e_gpu <= 3.0 * e_o32
