"""Pins the FP64 torch.fft autograd reference (oracle/torch_ref.py) on the CPU:
its forward map equals the oracle's LS-IMEX-RK3 trajectory, and its autograd
gradient equals central finite differences."""
import numpy as np
import pytest
import torch

from oracle import dns_oracle as O
from oracle import torch_ref as R


def exact_start(N, K0, seed):
    S0 = O.create_state(N, 1e3, K0, 0.5, seed, "PAO", dtype=np.float64)
    H = N // 2
    om = 1j * (np.array([z if z < H else z - N for z in range(N)])[:, None] * S0.uc0[0].astype(complex)
               - np.arange(H)[None, :] * S0.uc0[1].astype(complex))
    return om


def oracle_traj(om0, N, dt, visc, nsteps):
    H = N // 2
    ax = np.arange(H)[None, :]
    gz = np.array([z if z < H else z - N for z in range(N)])[:, None]
    k2 = ax * ax + gz * gz
    psi = np.where(k2 > 0, om0 / np.where(k2 > 0, k2, 1), 0)
    uc0 = np.stack([-1j * gz * psi, 1j * ax * psi])
    uc0[0][:, 0] = np.where(gz[:, 0] != 0, -1j * om0[:, 0] / np.where(gz[:, 0] != 0, gz[:, 0], 1), 0)
    S = O.create_state(N, 1e3, 2.0, 0.5, 1, "PAO", dtype=np.float64, uc0=uc0)
    S.arrays["om2"] = om0.copy()
    S.visc = visc
    O.step2a(S)
    S.dt = dt
    for _ in range(nsteps):
        O.rk3(S)
    return S.om2


def test_torch_reference_forward_equals_oracle():
    N, dt, visc, nsteps = 32, 2e-3, 1e-3, 6
    om0 = exact_start(N, 3.0, 5)
    geo = R.Geometry(N)
    omT = R.rk3_trajectory(torch.tensor(om0), geo, dt, visc, nsteps).numpy()
    ref = oracle_traj(om0, N, dt, visc, nsteps)
    assert np.linalg.norm(omT - ref) <= 1e-12 * np.linalg.norm(ref)


def test_torch_reference_gradient_matches_central_differences():
    N, dt, visc, nsteps = 16, 5e-3, 1e-3, 4
    om0 = exact_start(N, 2.0, 7)
    target = exact_start(N, 2.0, 8)
    L, g, _ = R.loss_and_grad(om0, target, N, dt, visc, nsteps)
    rng = np.random.default_rng(1)
    geo = R.Geometry(N)
    w = np.full((N, N // 2), 2.0)
    w[:, 0] = 1.0

    def loss(x):
        om = R.rk3_trajectory(torch.tensor(x), geo, dt, visc, nsteps).numpy()
        d = om - target
        return 0.5 * float(np.sum(w * (d.real ** 2 + d.imag ** 2)))
    for _ in range(3):
        d = rng.standard_normal(om0.shape) + 1j * rng.standard_normal(om0.shape)
        eps = 1e-5 * np.linalg.norm(om0) / np.linalg.norm(d)
        fd = (loss(om0 + eps * d) - loss(om0 - eps * d)) / (2 * eps)
        an = float(np.sum((np.conj(g) * d).real))
        assert abs(fd - an) <= 1e-7 * abs(an)
