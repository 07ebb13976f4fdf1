"""FP64 torch.fft autograd reference for the adjoint -- TEST INFRASTRUCTURE ONLY.

Restates the fixed-dt LS-IMEX-RK3 step of the reference
(turbo_simulator.py:1900-1943 with 1378-1513, 1520-1585, 1949-2037) in
differentiable torch on the CPU, with the velocity reconstructed from omega
at every stage, so that autograd gives an independent gradient of the
final-time vorticity mismatch with respect to omega_0.
"""
import math

import numpy as np
import torch

RK3_ALPHA = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)
RK3_BETA = (0.0, -17.0 / 60.0, -5.0 / 12.0)


class Geometry:
    def __init__(self, N):
        self.N, self.H, self.M = N, N // 2, 3 * N // 2
        ax = torch.arange(self.H, dtype=torch.float64)[None, :]
        gz = torch.tensor([z if z < self.H else z - N for z in range(N)], dtype=torch.float64)[:, None]
        self.ax, self.gz = ax, gz
        self.K2 = ax * ax + gz * gz
        self.GA = gz * ax
        self.G2mA2 = gz * gz - ax * ax
        self.divxz = 1.0 / ((1.5 * N) * (1.5 * N))
        invK2 = 1.0 / (self.K2 + 1e-30)
        invg = torch.where((torch.arange(N)[:, None] >= 1) & (gz != 0), 1.0 / torch.where(gz != 0, gz, 1.0), 0.0)
        cu = torch.where(ax >= 1, -1j * gz * invK2, -1j * invg)
        cw = torch.where(ax >= 1, 1j * ax * invK2, 0.0 + 0j)
        self.cu, self.cw = cu.to(torch.complex128), cw.to(torch.complex128)
        rows = torch.tensor([z if z < self.H else z + self.H for z in range(N)])
        self.rows = rows
        self.keep = torch.ones(N, 1, dtype=torch.float64)
        self.keep[self.H] = 0.0

    def pad(self, band):
        """compact (N, H) -> padded (M, M/2+1) spectrum (turbo_simulator.py:2020-2034)."""
        out = torch.zeros(self.M, self.M // 2 + 1, dtype=torch.complex128)
        return out.index_put((self.rows[:, None], torch.arange(self.H)[None, :]), band)

    def nonlinear(self, om):
        u = torch.fft.irfft2(self.pad(self.cu * om), s=(self.M, self.M), norm="forward")
        w = torch.fft.irfft2(self.pad(self.cw * om), s=(self.M, self.M), norm="forward")
        A = torch.fft.rfft2(u * u - w * w)[self.rows][:, : self.H]
        C = torch.fft.rfft2(u * w)[self.rows][:, : self.H]
        return (self.GA * A + self.G2mA2 * C) * self.divxz * self.keep


def rk3_trajectory(om0, geo, dt, visc, nsteps):
    om = om0
    fnm1 = torch.zeros_like(om0)
    for _ in range(nsteps):
        for a, b in zip(RK3_ALPHA, RK3_BETA):
            fn = geo.nonlinear(om)
            tmp = 1.0 - geo.K2 * (b * dt * visc)
            den = geo.K2 * (a * dt * visc) + 1.0
            om = (om * tmp + fn * (a * dt) + fnm1 * (b * dt)) / den
            fnm1 = fn
    return om


def loss_and_grad(om0_np, target_np, N, dt, visc, nsteps):
    geo = Geometry(N)
    x = torch.tensor(om0_np, dtype=torch.complex128, requires_grad=True)
    omT = rk3_trajectory(x, geo, dt, visc, nsteps)
    w = torch.full((N, N // 2), 2.0, dtype=torch.float64)
    w[:, 0] = 1.0
    d = omT - torch.as_tensor(target_np, dtype=torch.complex128)
    L = 0.5 * (w * (d.real ** 2 + d.imag ** 2)).sum()
    L.backward()
    return float(L), x.grad.numpy(), omT.detach().numpy()
