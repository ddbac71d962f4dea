"""Waveform misfit as a torch.autograd.Function.

forward(v) runs the GPU forward pass, forms the residual against the
observed traces and the misfit J = 1/2 sum (d - d_obs)^2.  When the velocity
needs a gradient, the same call also runs the hand-written reverse pass
(stp_fd_adjoint: checkpointed forward recompute + adjoint recursion) and keeps
dJ/dv for backward; otherwise (line searches, validation, torch.no_grad) it
costs one forward pass.  Velocity is the interior (nx, ny, nz) model;
gradients are FP32 on the device.
"""
from __future__ import annotations

import torch

from .fd import FDPropagator


class WaveformMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, velocity: torch.Tensor, prop: FDPropagator, wavelet, obs: torch.Tensor, nt: int,
                checkpoint_interval: int = 0):
        prop.set_velocity(velocity.detach())
        if ctx.needs_input_grad[0]:
            J, grad, traces = prop.misfit_gradient(nt, wavelet, obs, checkpoint_interval)
            ctx.save_for_backward(grad)
        else:
            J, traces = prop.misfit(nt, wavelet, obs)
        ctx.mark_non_differentiable(traces)
        return torch.tensor(J, dtype=torch.float64, device=velocity.device), traces

    @staticmethod
    def backward(ctx, grad_j, grad_traces):
        (grad,) = ctx.saved_tensors
        return (grad_j.to(grad.dtype) * grad).to(torch.float32), None, None, None, None, None


def waveform_misfit(velocity, prop: FDPropagator, wavelet, obs, nt: int, checkpoint_interval: int = 0):
    """J(velocity) (float64 scalar tensor) and the modelled traces; J.backward()
    gives velocity.grad = dJ/dv.  Without a gradient requested (velocity not
    requiring grad, or under torch.no_grad) only the forward pass runs."""
    return WaveformMisfit.apply(velocity, prop, wavelet, obs, nt, checkpoint_interval)
