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


def _to_handle(h, field, t: torch.Tensor):
    """Copy a tensor into a handle field: device to device for CUDA tensors
    (palx_upload copies from device pointers too), through NumPy otherwise."""
    if t.is_cuda:
        c = t.detach().to(dtype=torch.complex64 if h.precision == 4 else torch.complex128).contiguous()
        torch.cuda.current_stream().synchronize()  # c is complete before libpalx copies it
        h.upload_ptr(field, c.data_ptr(), c.numel() * c.element_size())
        return
    h.upload(field, t.detach().numpy())


def _from_handle(h, field, like: torch.Tensor) -> torch.Tensor:
    """A handle field as a tensor on like's device (no host copy for CUDA)."""
    shp, dt = h.shape(field)
    if like.is_cuda:
        out = torch.empty(shp, dtype=torch.complex64 if dt == np.complex64 else torch.complex128,
                          device=like.device)
        h.download_ptr(field, out.data_ptr(), out.numel() * out.element_size())
        return out
    return torch.from_numpy(h.download(field))


class RK3Trajectory(torch.autograd.Function):
    """omega_T = Phi(omega_0) on one libpalx handle (dt, visc already set).

    CUDA tensors stay on the device (no host round trip).  The handle keeps
    one trajectory's checkpoints: running another forward on the same handle
    before this one's backward makes that backward raise instead of
    returning the other trajectory's gradient."""

    @staticmethod
    def forward(ctx, om0, handle, nsteps, ckpt_every):
        h = handle
        _to_handle(h, P.F_OM, om0)
        h.adj_forward(nsteps, ckpt_every)
        h.adj_generation = getattr(h, "adj_generation", 0) + 1
        ctx.handle = h
        ctx.generation = h.adj_generation
        return _from_handle(h, P.F_OM, om0)

    @staticmethod
    def backward(ctx, grad_omT):
        h = ctx.handle
        if getattr(h, "adj_generation", 0) != ctx.generation:
            raise RuntimeError("another RK3Trajectory forward ran on this handle after this one: its "
                               "checkpoints are gone; use one handle per live trajectory")
        if grad_omT.is_cuda:  # lambda_T read straight from the device tensor ([z][kx], contiguous)
            g = grad_omT.detach().to(dtype=torch.complex64 if h.precision == 4 else torch.complex128).contiguous()
            torch.cuda.current_stream().synchronize()
            h.call("palx_adj_backward", P.C.c_void_p(g.data_ptr()))
        else:
            h.adj_backward(grad_omT.detach().numpy())
        return _from_handle(h, P.F_ADJ_OM, grad_omT), None, None, None


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
