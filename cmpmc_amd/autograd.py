"""torch.autograd.Function over the differentiable batched transport.

    y = transport_observables(theta, engine, workload, n_events, seed, t_max_s)

theta is a float64 tensor (B_iso, A_anh, f_diffuse, p_film); y (66) holds the energy
collected on the top and bottom faces and the energy-weighted 64-bin arrival-time
histogram over [0, t_max_s), as batch totals. forward() runs the HIP transport kernel
and keeps its hit tape; backward() runs the HIP backward kernel on that tape with the
incoming gradient. The gradient is a likelihood-ratio (score) estimate of
d E[y] / d theta: unbiased, with Monte-Carlo noise that falls as 1 / sqrt(n_events).
"""
from __future__ import annotations

import numpy as np
import torch

from . import Engine
from ._capi import Workload


class TransportObservables(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta: torch.Tensor, engine: Engine, wl: Workload, n_events: int, seed: int,
                t_max_s: float) -> torch.Tensor:
        params = theta.detach().to("cpu", torch.float64).numpy()
        y, tape, info = engine.diff_forward(wl, n_events, params, seed=seed, t_max_s=t_max_s)
        ctx.tape = tape
        ctx.info = info
        return torch.from_numpy(y).to(device=theta.device)

    @staticmethod
    def backward(ctx, gy: torch.Tensor):
        g = gy.detach().to("cpu", torch.float64).numpy()
        grad = ctx.tape.backward(np.ascontiguousarray(g))
        return torch.from_numpy(grad).to(device=gy.device), None, None, None, None, None


def transport_observables(theta: torch.Tensor, engine: Engine, wl: Workload, n_events: int, seed: int = 1,
                          t_max_s: float = 1e-3) -> torch.Tensor:
    return TransportObservables.apply(theta, engine, wl, n_events, seed, t_max_s)
