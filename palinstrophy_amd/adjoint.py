"""Gradient of a final-time vorticity mismatch with respect to omega_0.

The forward map is the fixed-dt LS-IMEX-RK3 trajectory of the DNS step
(turbo_simulator.py:1900-1943) with the velocity always reconstructed from
omega (turbo_simulator.py:1555-1585); the reference has no adjoint.  The
backward pass is the hand-written discrete adjoint in libpalx
(palx_adj_forward / palx_adj_backward): checkpoints of (omega, fnm1) every
``ckpt_every`` steps, each segment replayed on the GPU and swept backwards
stage by stage.  Gradients follow torch's convention for complex leaves:
dL/dRe + i dL/dIm of every stored mode of the compact (N, N/2) spectrum.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _palx as P


def half_spectrum_weights(N: int) -> np.ndarray:
    """Real inner product of the half spectrum: 1 on the kx=0 column, 2 elsewhere."""
    w = np.full((N, N // 2), 2.0)
    w[:, 0] = 1.0
    return w


class RK3Trajectory(torch.autograd.Function):
    """omega_T = Phi(omega_0) on one libpalx handle (dt, visc already set)."""

    @staticmethod
    def forward(ctx, om0, handle, nsteps, ckpt_every):
        h = handle
        h.upload(P.F_OM, om0.detach().cpu().numpy())
        h.adj_forward(nsteps, ckpt_every)
        omT = h.download(P.F_OM)
        ctx.handle = h
        return torch.from_numpy(omT).to(om0.device)

    @staticmethod
    def backward(ctx, grad_omT):
        g = ctx.handle.adj_backward(grad_omT.detach().cpu().numpy())
        return torch.from_numpy(g).to(grad_omT.device), None, None, None


def mismatch_loss(omT: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """L = 1/2 || omega_T - target ||^2 with the half-spectrum weights."""
    w = torch.as_tensor(half_spectrum_weights(omT.shape[0]), dtype=omT.real.dtype, device=omT.device)
    d = omT - target
    return 0.5 * (w * (d.real ** 2 + d.imag ** 2)).sum()


def loss_and_grad(handle, om0: np.ndarray, target: np.ndarray, nsteps: int, ckpt_every: int = 20):
    """(loss, dL/d omega_0) for the mismatch loss, via the autograd Function."""
    x = torch.tensor(om0, requires_grad=True)
    omT = RK3Trajectory.apply(x, handle, nsteps, ckpt_every)
    L = mismatch_loss(omT, torch.as_tensor(target))
    L.backward()
    return float(L), x.grad.numpy()
