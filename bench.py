#!/usr/bin/env python3
"""Benchmark of the palinstrophy DNS time step on one MI355X.

Workload (the configuration the headline metric is quoted on): N=4096
(6144^2 padded FFTs), KM3 spectrum, K0=15, Re from the repo's Re(N,K0) rule
(1.094843e5), seed 4895, LS-IMEX-RK3 at CFL 0.5 with the reference NEXTDT
cadence (it==1 or it%100==0).  The bulk FFT / nonlinear arithmetic runs in
FP32 (t, dt, CFL and diagnostic sums in FP64); the same workload in FP64 is
reported beside it.  A "step" is one full RK3 step (3 stages) of the whole
field; the metric is grid-point-steps/s = N^2 * steps/s on the base grid.

Per-kernel durations come from HIP events recorded on the handle's stream
(torch's current stream) around every kernel launch of every 4th step of the
timed region.
"""
from __future__ import annotations

import argparse
import contextlib
import json
import math
import os
import sys
import time

import numpy as np

HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s

HERE = os.path.dirname(os.path.abspath(__file__))


def re_from_n_k0(N, K0):
    # turbo_main.py:2323-2328
    return 10.0 ** (1.066240 * math.log10(N) - 0.083911 * math.log10(K0) + 1.286396)


def algo_bytes(N, prec, method="LS_IMEX_RK3"):
    """Compulsory HBM bytes of one stage = (row pass, column pass), the
    column pass averaged over the stages of a step.

    row:    read + write 2 planes of M x N/2 complex            = 6  s N^2
    column: read + write 2 planes of M x N/2 complex (6 s N^2) and
            N x N/2 complex planes of omega / fnm1 (s N^2 each):
            CNAB2 reads and writes omega and fnm1                = 10 s N^2;
            RK3, stages of one step (STAGE_IN_STEP): stage 0 reads
            omega and writes the next stage's pre-update term, stage 1
            reads and writes that term, stage 2 reads it and writes
            omega and fnm1 -- 7 planes in 3 stages               = 25/3 s N^2
            per stage (three plain stages move 11: 29/3 s N^2)
    (s = bytes per real; M = 3N/2).
    """
    s = prec
    M, H = 3 * N // 2, N // 2
    c = 2 * s
    row = 4 * M * H * c
    if method == "LS_IMEX_RK3":
        col = 4 * M * H * c + (7 * N * H * c) // 3
    else:
        col = (4 * M * H + 4 * N * H) * c
    return row, col


def run_gpu(N, prec, steps, warmup, seed=4895, K0=15.0, cfl=0.5, update=100, tune=True, variants=None,
            method="LS_IMEX_RK3"):
    import torch
    from palinstrophy_amd import _palx as P
    from palinstrophy_amd import autotune

    torch.cuda.init()
    stream = torch.cuda.current_stream()
    Re = re_from_n_k0(N, K0)
    uc0, stats = P.pao_generate(N, K0, Re, seed, True)
    om0 = P.calcom_f32(uc0)
    h = P.Handle(N, prec)
    tuning = None
    if variants is not None:
        h.set_variants(*variants)
    elif tune:  # radix / threads / LDS layout of both passes, timed on a scratch field
        tuning = autotune.tune_handle(h, reps=10)
    h.call("palx_set_stream", stream.cuda_stream)
    h.load_initial(uc0, om0)
    h.set_scalar(P.S_VISC, stats[4])
    h.set_scalar(P.S_CFLNUM, cfl)
    cflm0 = h.cflm()
    h.set_scalar(P.S_DT, cfl / (cflm0 * math.pi))
    h.set_scalar(P.S_CN, 1.0)
    h.set_scalar(P.S_CNM1, 0.0)
    h.set_scalar(P.S_T, 0.0)

    mid = P.CNAB2 if method == "CNAB2" else P.LS_IMEX_RK3
    nst = 1 if method == "CNAB2" else 3
    # the RK3 stages of a step as dns_step_ls_imex_rk3 runs them (STAGE_IN_STEP)
    flag = 0 if method == "CNAB2" else P.STAGE_IN_STEP

    def one_step(it, events=None):
        for st in range(nst):
            if events is not None:
                e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
                e0.record(stream)
                h.stage_products()
                e1.record(stream)
                h.stage_update(mid, st | flag)
                e2.record(stream)
                events.append((e0, e1, e2))
            else:
                h.stage_products()
                h.stage_update(mid, st | flag)
        if it == 1 or it % update == 0:
            h.next_dt()

    it = 0
    for _ in range(warmup):
        it += 1
        one_step(it)
    torch.cuda.synchronize()
    # HIP events around every launch of every 4th timed step: events between
    # all launches cost 2.6% of the step on MI355X, every 4th step 0.8%
    events = []
    t0 = time.perf_counter()
    for k in range(steps):
        it += 1
        one_step(it, events if k % 4 == 0 else None)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    row_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in events]))
    col_ms = float(np.mean([b.elapsed_time(c) for _, b, c in events]))
    # NEXTDT (physical pass + CFL reduction + dt update) on its own: it runs
    # once every `update` steps, so a short timed window may hold none
    n0, n1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    n0.record(stream)
    for _ in range(5):
        h.next_dt()
    n1.record(stream)
    torch.cuda.synchronize()
    nextdt_ms = n0.elapsed_time(n1) / 5
    sc = h.scalars()
    ok = bool(np.isfinite(sc[P.S_DT]) and sc[P.S_NAN] == 0.0)
    vr, vc = h.variants()
    h.close()
    nextdt_in_window = sum(1 for k in range(warmup + 1, warmup + steps + 1) if k == 1 or k % update == 0)
    return dict(elapsed=elapsed, steps=steps, row_ms=row_ms, col_ms=col_ms, ok=ok,
                dt=float(sc[P.S_DT]), cflm=float(sc[P.S_CFLM]), variants=(vr, vc), tuning=tuning,
                nextdt_ms=nextdt_ms, nextdt_in_window=nextdt_in_window)


def precision_errors(h32, h64, nbins):
    """FP32 path against the FP64 path: omega relative L2, energy spectrum
    (relative L2 over bins and max relative error over bins holding more
    than 1e-12 of the peak) and enstrophy."""
    from palinstrophy_amd import _palx as P

    om32 = h32.download(P.F_OM).astype(np.complex128)
    om64 = h64.download(P.F_OM)
    e_om = float(np.linalg.norm(om32 - om64) / np.linalg.norm(om64))
    s32, _ = h32.spectrum(nbins)
    s64, _ = h64.spectrum(nbins)
    sig = s64 > 1e-12 * s64.max()
    e_spec = float(np.linalg.norm(s32 - s64) / np.linalg.norm(s64))
    e_bin = float(np.max(np.abs(s32[sig] - s64[sig]) / s64[sig]))
    z32, z64 = h32.diag()[3], h64.diag()[3]
    return {"omega_rel_l2": e_om, "spectrum_rel_l2": e_spec, "spectrum_max_bin_rel": e_bin,
            "enstrophy_rel": float(abs(z32 - z64) / abs(z64)),
            "dt32": float(h32.scalars()[P.S_DT]), "dt64": float(h64.scalars()[P.S_DT])}


def precision_run(N=4096, checkpoints=(10, 100, 500), seed=4895, K0=15.0, cfl=0.5, update=100):
    """Precision form: the same seeded KM3 field (K0=15, Re from the Re(N,K0)
    rule) advanced with LS-IMEX-RK3 at CFL 0.5 by the FP32 and the FP64 path,
    each with its own adaptive dt (NEXTDT it==1 or it%100==0); errors of the
    FP32 path against the FP64 path at the given steps (nbins = 2N)."""
    from palinstrophy_amd import _palx as P
    from palinstrophy_amd import autotune

    Re = re_from_n_k0(N, K0)
    uc0, stats = P.pao_generate(N, K0, Re, seed, True)
    om0 = P.calcom_f32(uc0)
    hs = {}
    for prec in (4, 8):
        h = P.Handle(N, prec)
        autotune.tune_handle(h, reps=5)
        h.load_initial(uc0, om0)
        h.set_scalar(P.S_VISC, stats[4])
        h.set_scalar(P.S_CFLNUM, cfl)
        h.set_scalar(P.S_CN, 1.0)
        h.set_scalar(P.S_DT, cfl / (h.cflm() * math.pi))
        hs[prec] = h
    out, done = {}, 0
    for k in checkpoints:
        for h in hs.values():
            h.step(k - done, P.LS_IMEX_RK3, update)
        done = k
        out[str(k)] = precision_errors(hs[4], hs[8], max(32, 2 * N))
        print(f"[bench] precision step {k}: omega {out[str(k)]['omega_rel_l2']:.2e} "
              f"E(k) {out[str(k)]['spectrum_rel_l2']:.2e} Z {out[str(k)]['enstrophy_rel']:.2e}",
              file=sys.stderr, flush=True)
    for h in hs.values():
        h.close()
    return {"N": N, "workload": f"KM3 K0={K0:g} seed {seed} Re={Re:.6e} RK3 CFL {cfl} adaptive", "steps": out}


SMALL_WORKLOADS = (
    # N, precision, method, CFL, spectrum, K0, Re, steps
    (1024, 8, "LS_IMEX_RK3", 0.1, "PAO", 15.0, 1.0e4, 1000),
    (256, 8, "CNAB2", 0.75, "KM3", 10.0, 1.0e4, 200),
)


def small_runs(seed=4895, update=100):
    """Small grids are launch-bound: the same adaptive run with eager
    palx_step launches and with one HIP graph launch per step
    (palx_set_graph).  N=256 CNAB2 at CFL 0.75 blows up after ~250 steps in the
    reference algorithm, so it is timed over 200."""
    from palinstrophy_amd import _palx as P
    from palinstrophy_amd.ensemble import DnsEnsemble, Member

    out = []
    for N, prec, method, cfl, spec, K0, Re, steps in SMALL_WORKLOADS:
        mid = P.LS_IMEX_RK3 if method == "LS_IMEX_RK3" else P.CNAB2
        m = Member(N, seed, K0, spec, Re)
        res = {"N": N, "dtype": "f64" if prec == 8 else "f32", "method": method, "CFL": cfl, "spectrum": spec,
               "K0": K0, "Re": Re, "steps": steps}
        for mode in ("eager", "graph"):
            ens = DnsEnsemble([m], precision=prec, CFL=cfl, tune=True, tune_reps=5)
            h = ens.handles[0]
            h.set_graph(mode == "graph")
            h.step(2, mid, update)
            h.sync()
            t0 = time.perf_counter()
            h.step(steps, mid, update)
            h.sync()
            el = time.perf_counter() - t0
            res[f"{mode}_ms_per_step"] = el / steps * 1e3
            res[f"{mode}_value"] = N * N * steps / el
            res["finite"] = bool(np.isfinite(h.scalars()[P.S_DT]))
            ens.close()
        out.append(res)
        print(f"[bench] N={N} {method} {res['dtype']}: eager {res['eager_ms_per_step']:.4f} ms/step, "
              f"graph {res['graph_ms_per_step']:.4f} ms/step", file=sys.stderr, flush=True)
    return out


def frame_run(N=4096, frames=10, seed=4895, K0=15.0, cfl=0.5):
    """Cost of one rendered display frame at N (the GUI frame of
    turbo_main.py:2075-2125), omega variable, display stretch, at the GUI's
    display scale:
      device: palx_frame (getter, map, statistics, metrics, controller,
              stretch, striding) and one copy of the display-sized buffer;
      host finish: the reference's split -- the full-resolution [1,255] plane
              copied to the host, then NumPy float32 mean / std / stretch /
              striding on the host (turbo_main.py:2080-2114)."""
    from palinstrophy_amd import _palx as P
    from palinstrophy_amd import turbo_simulator as ts
    from palinstrophy_amd.frame import FrameBuffer, display_scale, frame_ctl, scale_arg

    with contextlib.redirect_stdout(sys.stderr):  # the init banner; stdout carries only the JSON line
        S = ts.create_dns_state(N, re_from_n_k0(N, K0), K0, cfl, "gpu", seed, start_spectrum="KM3",
                                precision="fp32")
    S.dt = cfl / (ts.compute_cflm(S) * math.pi)
    S.handle.step(20, P.LS_IMEX_RK3, 100)
    ds = display_scale(N)
    dev = FrameBuffer(S, scale_arg(ds))
    raw = FrameBuffer(S, 1)
    ctl = frame_ctl(N, K0)
    S.handle.frame_reset(S.Re)
    flags = P.FRAME_STRETCH | P.FRAME_ADAPT
    dev.run(P.VAR_OMEGA, flags, ctl)
    t0 = time.perf_counter()
    for _ in range(frames):
        fr = dev.run(P.VAR_OMEGA, flags, ctl)
    t_dev = (time.perf_counter() - t0) / frames
    s = int(ds)

    def host_finish():
        pixels = raw.run(P.VAR_OMEGA, 0, None).pixels
        pix_f = pixels.astype(np.float32, copy=False)
        mu, sig = float(pix_f.mean()), float(pix_f.std())
        lo, hi = mu - 2.5 * sig, mu + 2.5 * sig
        inv = 255.0 / (hi - lo) if (hi - lo) != 0.0 else 0.0
        out = ((pix_f - lo) * inv).round().clip(0.0, 255.0).astype(np.uint8)
        return np.ascontiguousarray(out[::s, ::s]) if s > 1 else out

    host_finish()
    hf = max(2, frames // 3)
    t0 = time.perf_counter()
    for _ in range(hf):
        out = host_finish()
    t_host = (time.perf_counter() - t0) / hf
    res = {"N": N, "display": list(fr.pixels.shape), "device_ms_per_frame": t_dev * 1e3,
           "host_finish_ms_per_frame": t_host * 1e3, "speedup": t_host / t_dev,
           "bytes_to_host_device": dev.nbytes, "bytes_to_host_host_finish": raw.nbytes,
           "same_display_size": list(out.shape) == list(fr.pixels.shape)}
    print(f"[bench] frame N={N}: device {res['device_ms_per_frame']:.3f} ms, host finish "
          f"{res['host_finish_ms_per_frame']:.1f} ms", file=sys.stderr, flush=True)
    return res


def ensemble_run(steps=200, warmup=5, prec=4, meta_seed=2024):
    """Batched ensemble workload: 64 runs at N=256, 16 at N=512 and 4 at
    N=1024 (seeds 1..84, K0 from {5,10,15}, KM3 or PAO), LS-IMEX-RK3 at CFL
    0.5, own adaptive dt per member (NEXTDT it==1 or it%100==0).  Aggregate
    grid-point-steps/s with one graph launch per ensemble step, against the
    same members stepped one after another on their own handles (same
    kernels and variants, palx_step)."""
    import torch
    from palinstrophy_amd import _palx as P
    from palinstrophy_amd.ensemble import DnsEnsemble, mixed_ensemble

    members = mixed_ensemble(meta_seed=meta_seed)
    modes = {}
    for mode in ("graph", "streams", "serial"):
        ens = DnsEnsemble(members, precision=prec, CFL=0.5, tune=True)
        gpts = ens.grid_points()
        ens.ens.step(warmup, P.LS_IMEX_RK3, 100, graph=(mode == "graph"), serial=(mode == "serial"))
        ens.sync()
        t0 = time.perf_counter()
        ens.ens.step(steps, P.LS_IMEX_RK3, 100, graph=(mode == "graph"), serial=(mode == "serial"))
        ens.sync()
        modes[mode] = time.perf_counter() - t0
        finite = all(np.isfinite(h.scalars()[P.S_DT]) for h in ens.handles)
        variants = {N: (vr, vc) for N, _, vr, vc in ens.ens.classes()}
        ens.close()
    t_graph = modes["graph"]
    # sequential: the same members, each on its own handle, one after another
    handles = []
    for m in members:
        h = P.Handle(m.N, prec)
        h.set_variants(*variants[m.N])
        uc0, st = P.pao_generate(m.N, m.K0, m.reynolds(), m.seed, m.spectrum == "KM3")
        h.load_initial(uc0, P.calcom_f32(uc0))
        h.set_scalar(P.S_VISC, st[4])
        h.set_scalar(P.S_CFLNUM, 0.5)
        h.set_scalar(P.S_CN, 1.0)
        h.set_scalar(P.S_DT, 0.5 / (h.cflm() * math.pi))
        h.step(warmup, P.LS_IMEX_RK3, 100)
        handles.append(h)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for h in handles:
        h.step(steps, P.LS_IMEX_RK3, 100)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    for h in handles:
        h.close()
    counts = {}
    for m in members:
        counts[m.N] = counts.get(m.N, 0) + 1
    return {"members": {str(k): v for k, v in sorted(counts.items())}, "steps": steps, "dtype": "f32" if prec == 4 else "f64",
            "grid_points": gpts, "value": gpts * steps / t_graph, "unit": "grid-pt-steps/s",
            "ms_per_step": t_graph / steps * 1e3, "sequential_value": gpts * steps / t_seq,
            "speedup_vs_sequential": t_seq / t_graph, "variants": {str(k): v for k, v in variants.items()},
            "ms_per_step_by_launch": {"graph": modes["graph"] / steps * 1e3,
                                      "forked_streams": modes["streams"] / steps * 1e3,
                                      "one_stream": modes["serial"] / steps * 1e3},
            "finite": finite}


def _cpu_rate(N, dtype, seconds_budget, seed, K0, cfl, workers):
    from oracle import dns_oracle as O
    from palinstrophy_amd import _palx as P

    Re = re_from_n_k0(N, K0)
    uc0, _ = P.pao_generate(N, K0, Re, seed, True)
    S = O.create_state(N, Re, K0, cfl, seed, "KM3", dtype=dtype, workers=workers, uc0=uc0)
    O.init_dt(S)
    t0 = time.perf_counter()
    O.rk3(S)  # warm-up (FFT planning)
    one = time.perf_counter() - t0
    k = max(1, min(40, int(seconds_budget / max(one, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(k):
        O.rk3(S)
    el = time.perf_counter() - t0
    return N * N * k / el, k, el


def cpu_baseline(N, seconds_budget=12.0, seed=4895, K0=15.0, cfl=0.5):
    """The oracle (NumPy + scipy.fft, all host cores) on the same workload,
    timed over a bounded number of RK3 steps: complex64 (the reference's
    precision, `value`) and complex128 beside it."""
    workers = len(os.sched_getaffinity(0))
    v32, k32, e32 = _cpu_rate(N, np.float32, seconds_budget, seed, K0, cfl, workers)
    v64, k64, e64 = _cpu_rate(N, np.float64, seconds_budget, seed, K0, cfl, workers)
    return dict(value=v32, unit="grid-pt-steps/s", cores=workers, kind="port",
                sample=(f"{k32} LS-IMEX-RK3 steps at N={N} (FP32 oracle, scipy.fft workers={workers}) "
                        f"after 1 warm-up step, {e32:.1f}s"),
                fp64={"value": v64, "sample": f"{k64} steps FP64 oracle, {e64:.1f}s"})


def adjoint_ratio(N=512, nsteps=200, ckpt=20, prec=4, tune=True):
    """Adjoint workload (N=512, Re=5000, PAO K0=10, seed 4895, 200 fixed-dt RK3 steps,
    checkpoints every 20): wall time of the backward sweep (segment replays
    included) over the checkpointed forward of the same trajectory.  With
    tune, the step passes (forward and replay) use the autotuner's variants;
    the adjoint passes always use the default ones."""
    import torch
    from palinstrophy_amd import _palx as P
    from palinstrophy_amd import autotune

    uc0, st = P.pao_generate(N, 10.0, 5000.0, 4895, False)
    om0 = P.calcom_f32(uc0)
    h = P.Handle(N, prec)
    if tune:
        autotune.tune_handle(h, reps=10)
    h.call("palx_set_stream", torch.cuda.current_stream().cuda_stream)
    h.set_scalar(P.S_VISC, st[4])
    h.upload(P.F_OM, om0)
    h.recon_velocity()
    h.set_scalar(P.S_DT, 0.3 / (h.cflm() * math.pi))
    lam = np.ones((N, N // 2), dtype=np.complex64 if prec == 4 else np.complex128)
    h.upload(P.F_OM, om0)
    h.adj_forward(nsteps, ckpt)
    h.adj_backward(lam)           # warm-up (allocations, code objects)
    torch.cuda.synchronize()
    h.upload(P.F_OM, om0)
    t0 = time.perf_counter()
    h.adj_forward(nsteps, ckpt)
    torch.cuda.synchronize()
    tf = time.perf_counter() - t0
    t0 = time.perf_counter()
    h.adj_backward(lam)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    vr, vc = h.variants()
    h.close()
    return {"N": N, "steps": nsteps, "ckpt_every": ckpt, "dtype": "f32" if prec == 4 else "f64",
            "forward_ms": tf * 1e3, "adjoint_ms": tb * 1e3, "adjoint_over_forward": tb / tf,
            "step_variants": [vr, vc]}


def traffic_sweep(N, K0=15.0, seed=4895, cfl=0.5):
    """One RK3 step (stage_products + stage_update for stages 0, 1, 2, in that
    order, as one step: STAGE_IN_STEP) of every compiled step-pass variant,
    FP32 then FP64, with pinned variants and no tuner: under a rocprof counter
    run each k_row / k_col variant then has exactly three launches, one per
    RK3 stage, whose mean is the step-average traffic of that kernel (the
    figure algo_bytes models)."""
    from palinstrophy_amd import _palx as P

    Re = re_from_n_k0(N, K0)
    uc0, stats = P.pao_generate(N, K0, Re, seed, True)
    om0 = P.calcom_f32(uc0)
    for prec in (4, 8):
        h = P.Handle(N, prec)
        h.load_initial(uc0, om0)
        h.set_scalar(P.S_VISC, stats[4])
        h.set_scalar(P.S_CFLNUM, cfl)
        h.set_scalar(P.S_DT, cfl / (h.cflm() * math.pi))
        for v in range(P.n_variants(N)):
            if not P.variant_ok(N, prec, v):
                continue
            h.set_variants(v, v)
            print(f"[bench] traffic sweep: {'fp32' if prec == 4 else 'fp64'} variant {v}", file=sys.stderr, flush=True)
            for st in range(3):
                h.stage_products()
                h.stage_update(P.LS_IMEX_RK3, st | P.STAGE_IN_STEP)
            h.sync()
        h.close()


def load_traffic():
    p = os.path.join(HERE, "profiles", "traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def _traffic_bytes(traffic, kname, N, algo):
    """HBM bytes per launch of `kname` from the committed rocprof counter runs
    of `bench.py --traffic-sweep` (FETCH_SIZE and WRITE_SIZE scaled by the
    copy-kernel calibration; the mean of the variant's three launches, one
    per RK3 stage, the step average that `algo` models), and where they come
    from.  (None, None) when no measurement exists for this kernel and size."""
    if not traffic or N != 4096:
        return None, None
    e = traffic.get("kernels", {}).get(kname)
    if not e:
        return None, None
    return e["bytes"], "rocprof FETCH_SIZE + WRITE_SIZE, one RK3 step of the variant (--traffic-sweep)"


def dominant_roofline(r, prec, N, traffic):
    """HBM roofline of the longer step pass of run `r` (RK3): its algorithmic
    bytes per launch over its mean HIP-event duration."""
    row_b, col_b = algo_bytes(N, prec)
    real = "float" if prec == 4 else "double"
    vr, vc = r["variants"]
    if r["col_ms"] >= r["row_ms"]:
        kname, kbytes, kms = "k_col<%s,%d,RK3,V%d>" % (real, 3 * N // 2, vc), col_b, r["col_ms"]
    else:
        kname, kbytes, kms = "k_row<%s,%d,STEP,V%d>" % (real, 3 * N // 2, vr), row_b, r["row_ms"]
    achieved = kbytes / (kms * 1e-3)
    tb, src = _traffic_bytes(traffic, kname, N, kbytes)
    return {"bound": "hbm", "kernel": kname, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": tb, "traffic_source": src,
            "algo_bytes_per_launch": kbytes}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--N", type=int, default=4096)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-small", action="store_true")
    ap.add_argument("--no-cnab2", action="store_true")
    ap.add_argument("--no-frame", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="default kernel variants instead of the autotuner's")
    ap.add_argument("--variants", default=None, metavar="ROW,COL",
                    help="pin the FP32 row / column pass variants (reproducible profiles)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run one plain device copy of MIB MiB (HBM counter calibration)")
    ap.add_argument("--traffic-sweep", action="store_true",
                    help="only run one RK3 step of every step-pass variant (HBM counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the DNS of one field on one "
              "MI355X (multi-GPU slab decomposition is a design note, not built); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        a = torch.ones(args.copy_calibration * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(f"[bench] copied {args.copy_calibration} MiB", file=sys.stderr)
        return
    if args.traffic_sweep:
        traffic_sweep(args.N)
        return
    N = args.N
    pinned = tuple(int(x) for x in args.variants.split(",")) if args.variants else None
    r32 = run_gpu(N, 4, args.steps, args.warmup, tune=not args.no_tune, variants=pinned)
    print(f"[bench] fp32 N={N}: {r32['elapsed'] / args.steps * 1e3:.3f} ms/step, "
          f"row {r32['row_ms']:.3f} ms, col {r32['col_ms']:.3f} ms", file=sys.stderr, flush=True)
    r64 = None
    if not args.no_fp64:
        r64 = run_gpu(N, 8, max(10, args.steps // 2), args.warmup, tune=not args.no_tune)
        print(f"[bench] fp64 N={N}: {r64['elapsed'] / r64['steps'] * 1e3:.3f} ms/step, "
              f"row {r64['row_ms']:.3f} ms, col {r64['col_ms']:.3f} ms", file=sys.stderr, flush=True)

    ms = r32["elapsed"] / args.steps * 1e3
    value = N * N * args.steps / r32["elapsed"]
    row_b, col_b = algo_bytes(N, 4)
    traffic = load_traffic()
    step_bytes = 3 * (row_b + col_b)
    out = {
        "metric": "grid-pt-steps/s at N=4096 (FP32 vs FP64 GPU, %HBM peak); adjoint/fwd ratio",
        "value": value,
        "unit": "grid-pt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded KM3 initial field (seed 4895), no dataset",
        "config": {"workload": f"DNS LS-IMEX-RK3 step, N={N} (padded {3 * N // 2}^2), KM3 K0=15, "
                               f"Re={re_from_n_k0(N, 15.0):.6e}, CFL 0.5 adaptive dt (NEXTDT it==1 or it%100==0)",
                   "N": N, "method": "LS_IMEX_RK3", "precision": "fp32 bulk, fp64 scalars"},
        "roofline": dominant_roofline(r32, 4, N, traffic),
        "hbm_frac_step": step_bytes * args.steps / r32["elapsed"] / HBM_PEAK,
        # the same with a reference-side byte model that carries three product
        # planes (F[u^2], F[w^2], F[uw]): 19 s N^2 per stage
        "hbm_frac_step_19sN2_model": 3 * 19 * 4 * N * N * args.steps / r32["elapsed"] / HBM_PEAK,
        "kernel_ms": {"row": r32["row_ms"], "col": r32["col_ms"]},
        # NEXTDT cost and how many ran in the timed window (the reference cadence
        # runs one per 100 steps: nextdt_ms / 100 per step)
        "nextdt": {"ms_per_call": r32["nextdt_ms"], "calls_in_timed_window": r32["nextdt_in_window"],
                   "ms_per_step_at_cadence": r32["nextdt_ms"] / 100.0},
        "variants": {"row": r32["variants"][0], "col": r32["variants"][1],
                     "tuning": r32["tuning"]},
        "ok": r32["ok"],
    }
    if r64 is not None:
        row64, col64 = algo_bytes(N, 8)
        out["fp64"] = {"value": N * N * r64["steps"] / r64["elapsed"],
                       "ms_per_step": r64["elapsed"] / r64["steps"] * 1e3,
                       "kernel_ms": {"row": r64["row_ms"], "col": r64["col_ms"]},
                       "variants": {"row": r64["variants"][0], "col": r64["variants"][1]},
                       "hbm_frac_step": 3 * (row64 + col64) * r64["steps"] / r64["elapsed"] / HBM_PEAK,
                       "roofline": dominant_roofline(r64, 8, N, traffic),
                       "ok": r64["ok"]}
        out["fp32_over_fp64"] = out["value"] / out["fp64"]["value"]
    if not args.no_cnab2:
        # the reference's published N=4096 figures predate LS-IMEX-RK3 (one stage per step)
        rc = run_gpu(N, 4, args.steps, args.warmup, tune=not args.no_tune, variants=pinned, method="CNAB2")
        out["cnab2"] = {"value": N * N * rc["steps"] / rc["elapsed"], "unit": "grid-pt-steps/s",
                        "ms_per_step": rc["elapsed"] / rc["steps"] * 1e3, "dtype": "f32",
                        "steps_per_s": rc["steps"] / rc["elapsed"],
                        "kernel_ms": {"row": rc["row_ms"], "col": rc["col_ms"]}, "ok": rc["ok"]}
        print(f"[bench] fp32 CNAB2 N={N}: {out['cnab2']['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    if not args.no_precision:
        try:
            out["precision"] = precision_run(N)
        except Exception as e:  # reported, not fatal for the headline line
            out["precision"] = {"error": str(e)}
    if not args.no_small:
        try:
            out["small_grids"] = small_runs()
        except Exception as e:  # reported, not fatal for the headline line
            out["small_grids"] = {"error": str(e)}
    if not args.no_frame:
        try:
            out["frame"] = frame_run(N)
        except Exception as e:  # reported, not fatal for the headline line
            out["frame"] = {"error": str(e)}
    if not args.no_ensemble:
        try:
            out["ensemble"] = ensemble_run()
        except Exception as e:  # reported, not fatal for the headline line
            out["ensemble"] = {"error": str(e)}
    if not args.no_adjoint:
        try:
            out["adjoint"] = adjoint_ratio()
        except Exception as e:  # reported, not fatal for the headline line
            out["adjoint"] = {"error": str(e)}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(N)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
