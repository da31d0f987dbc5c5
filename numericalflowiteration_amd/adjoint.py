"""torch.autograd wrapper of the device adjoint (nufi_adjoint_grad).

`time_integrated_energy(a, conf)` returns J = sum_{n=0}^{Nt} dt * energy_n of a
1D run whose initial condition has the parameters `a` (the multimode
amplitudes a_1..a_n, or alpha for the Landau / two-stream families) as a
differentiable torch scalar: the forward pass runs the whole time loop on the
GPU, the backward pass returns the gradient from the reverse sweep that was
computed alongside it.  The tensors stay on the host; the device work goes
through libnufi_hip.so.
"""
from __future__ import annotations

import torch

from .config import F0_MULTIMODE_1D, FP64, Config
from .scheduler import GpuScheduler


def _with_params(conf: Config, values) -> Config:
    c = conf.copy()
    vals = [float(v) for v in values]
    if c.f0.kind == F0_MULTIMODE_1D:
        c.f0.n_modes = len(vals)
        for i, v in enumerate(vals):
            c.f0.a[i] = v
    else:
        if len(vals) != 1:
            raise ValueError("this f0 family has one parameter (alpha)")
        c.f0.alpha = vals[0]
    return c


def energy_and_grad(conf: Config, values, device: int = 0, precision: int = FP64):
    """(J, dJ/da) for the parameters `values` of conf's f0 family, with the trace
    of the given precision (FP64 or FP32C) in the forward and reverse sweeps."""
    with GpuScheduler(_with_params(conf, values), device=device, precision=precision) as s:
        return s.adjoint_grad()


class TimeIntegratedEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a: torch.Tensor, conf: Config, precision: int = FP64):
        J, g = energy_and_grad(conf, a.detach().cpu().double().tolist(), precision=precision)
        ctx.save_for_backward(torch.as_tensor(g, dtype=a.dtype))
        return torch.as_tensor(J, dtype=a.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None, None


def time_integrated_energy(a: torch.Tensor, conf: Config, precision: int = FP64) -> torch.Tensor:
    return TimeIntegratedEnergy.apply(a, conf, precision)
