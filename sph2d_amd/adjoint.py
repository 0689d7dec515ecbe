"""Gradients of gauge-elevation losses with respect to the piston motion,
through the hand-written adjoint of the RRSPH_HIP step.

    eta = PistonGauges.apply(a, b, run)        # torch tensors a, b (harmonics)
    loss = ((eta - target) ** 2).sum()
    loss.backward()                              # a.grad, b.grad from the adjoint

`run` is a `DifferentiableRun`: the context, the initial state, the number
of steps, the checkpoint interval and the gauge definition.  The forward
re-uploads the initial state, sets the Fourier piston
x(t) = sum_k a_k sin(k w t) + b_k cos(k w t) and runs the steps keeping a
checkpoint every `every` steps; the backward replays each segment from its
checkpoint and runs the step adjoints in reverse (rrsph_hip_adjoint_*).
Tensors live on the host: the coefficients and gauge histories are small,
the particle state never leaves the GPU.
"""
from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np

from . import hip as H


@dataclasses.dataclass
class DifferentiableRun:
    sim: "H.HipStep"
    state: dict                 # r, v, rho, itype of the initial state
    nsteps: int
    every: int
    spec: "H.GaugeSpec"
    omega: float

    def forward(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        self.sim.upload(self.state["r"], self.state["v"], self.state["rho"], self.state["itype"])
        self.sim.set_piston(self.omega, a, b)
        return self.sim.adjoint_forward(self.nsteps, self.every, self.spec)

    def backward(self, eta_bar: np.ndarray, ncoef: int) -> dict:
        """Gradients for eta_bar = dL/d eta: the ncoef coefficients of a and
        of b (every paddle's, paddle-major) and the initial state."""
        return self.sim.adjoint_backward(eta_bar, ncoef)


def _torch():
    import torch

    return torch


class PistonGauges:
    """torch.autograd.Function: (a, b) -> gauge histories (nsteps, ngauges)."""

    _fn = None

    @classmethod
    def apply(cls, a, b, run: DifferentiableRun):
        if cls._fn is None:
            torch = _torch()

            class _Fn(torch.autograd.Function):
                @staticmethod
                def forward(ctx, a_t, b_t, run_obj):
                    a_np = a_t.detach().double().cpu().numpy().copy()
                    b_np = b_t.detach().double().cpu().numpy().copy()
                    eta = run_obj.forward(a_np, b_np)
                    ctx.run = run_obj
                    ctx.coefs = (a_np, b_np)
                    ctx.tape = run_obj.sim.tape_generation
                    ctx.ncoef = a_t.numel()
                    ctx.dtype, ctx.device = a_t.dtype, a_t.device
                    return torch.from_numpy(eta).to(dtype=torch.float64, device=a_t.device)

                @staticmethod
                def backward(ctx, eta_bar):
                    # another forward on the same context has replaced this
                    # call's tape since: record this call's trajectory again
                    if ctx.run.sim.tape_generation != ctx.tape:
                        ctx.run.forward(*ctx.coefs)
                    g = ctx.run.backward(eta_bar.detach().double().cpu().numpy(), ctx.ncoef)
                    ga = torch.from_numpy(g["a"]).to(dtype=ctx.dtype, device=ctx.device)
                    gb = torch.from_numpy(g["b"]).to(dtype=ctx.dtype, device=ctx.device)
                    return ga, gb, None

            cls._fn = _Fn
        return cls._fn.apply(a, b, run)


def central_difference(run: DifferentiableRun, a: np.ndarray, b: np.ndarray, loss, which: str, k: int,
                       eps: float) -> float:
    """(loss(theta + eps e_k) - loss(theta - eps e_k)) / (2 eps) for one
    coefficient; loss maps the gauge history to a float."""
    vals = []
    for sgn in (1.0, -1.0):
        aa, bb = np.array(a, dtype=np.float64), np.array(b, dtype=np.float64)
        (aa if which == "a" else bb)[k] += sgn * eps
        vals.append(loss(run.forward(aa, bb)))
    return (vals[0] - vals[1]) / (2 * eps)
