"""Ensembles of independent DNS runs of uneven grid size on one GPU.

Each member is a full run of the reference's `sim` loop (run_dns,
turbo_simulator.py:2744-2859): its own seeded PAO/KM3 field
(_pao_build_ur_and_stats, 213-485), its own viscosity, CFL-adaptive dt and
NEXTDT cadence (next_dt, 2127-2167).  Members of equal N share one batched
launch per pass; every step of the whole ensemble is one HIP graph launch.
The reference runs one DnsState per process and has no ensemble form.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import _palx as P
from . import autotune


def re_from_n_k0(N: int, K0: float) -> float:
    """Re(N, K0) rule of the reference GUI (turbo_main.py:2323-2328)."""
    return 10.0 ** (1.066240 * math.log10(N) - 0.083911 * math.log10(K0) + 1.286396)


@dataclass
class Member:
    N: int
    seed: int
    K0: float
    spectrum: str  # "KM3" or "PAO"
    Re: float | None = None

    def reynolds(self) -> float:
        return float(self.Re) if self.Re is not None else re_from_n_k0(self.N, self.K0)


def mixed_ensemble(counts=((256, 64), (512, 16), (1024, 4)), meta_seed: int = 2024) -> list[Member]:
    """Members with seeds 1..n, K0 drawn uniformly from {5, 10, 15} and the
    spectrum from {KM3, PAO} by a generator seeded with `meta_seed`."""
    rng = np.random.default_rng(meta_seed)
    out, seed = [], 1
    for N, n in counts:
        for _ in range(n):
            K0 = float(rng.choice([5.0, 10.0, 15.0]))
            spec = "KM3" if rng.integers(0, 2) == 0 else "PAO"
            out.append(Member(int(N), seed, K0, spec))
            seed += 1
    return out


class DnsEnsemble:
    """A set of members advanced together.

    precision: 4 (FP32 bulk arithmetic, FP64 scalars) or 8.
    tune: pick the FFT pass variants per size class with the autotuner
    (cached), before the fields are loaded."""

    def __init__(self, members, precision: int = 4, CFL: float = 0.5, tune: bool = True,
                 tune_reps: int = 10, stream=None):
        self.members = list(members)
        self.precision = int(precision)
        self.CFL = float(CFL)
        self.ens = P.Ensemble([m.N for m in self.members], self.precision)
        if stream is not None:
            self.ens.call("palx_ens_set_stream", stream)
        self.tuning = autotune.tune_ensemble(self.ens, reps=tune_reps) if tune else None
        self.it = 0
        for m, h in zip(self.members, self.ens.members):
            self._load(m, h)

    def _load(self, m: Member, h: "P.Handle"):
        uc0, stats = P.pao_generate(m.N, m.K0, m.reynolds(), m.seed, m.spectrum == "KM3")
        h.load_initial(uc0, P.calcom_f32(uc0))
        h.set_scalar(P.S_VISC, stats[4])
        h.set_scalar(P.S_CFLNUM, self.CFL)
        h.set_scalar(P.S_CN, 1.0)
        h.set_scalar(P.S_CNM1, 0.0)
        h.set_scalar(P.S_T, 0.0)
        cflm = h.cflm()
        h.set_scalar(P.S_DT, self.CFL / (cflm * math.pi))  # turbo_simulator.py:2812-2823

    @property
    def handles(self):
        return self.ens.members

    def step(self, nsteps: int, method: int = P.LS_IMEX_RK3, update: int = 100, graph: bool = True):
        """Advance every member by nsteps steps; NEXTDT after step it when
        it == 1 or it % update == 0 (turbo_simulator.py:2842)."""
        self.ens.step(nsteps, method, update, graph=graph)
        self.it += nsteps

    def sync(self):
        self.ens.sync()

    def grid_points(self) -> int:
        return sum(m.N * m.N for m in self.members)

    def close(self):
        self.ens.close()
