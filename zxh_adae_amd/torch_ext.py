"""torch.autograd.Function over the engine's adjoint gradient.

PyTorch is used only here, as the autograd front end: the forward and the
backward both run in the native library (zxh_ansatz_forward / _backward).

    a = Ansatz(n, ops, nparams, "c64"); a.set_observable(terms)
    theta = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    e = expectation(theta, a)
    e.backward()            # theta.grad = dE/dtheta (adjoint method)
"""
from __future__ import annotations

import torch

from .ansatz import Ansatz


class Expectation(torch.autograd.Function):
    """E(theta) = <psi(theta)|H|psi(theta)> for an Ansatz with its observable."""

    @staticmethod
    def forward(ctx, params: torch.Tensor, ansatz: Ansatz) -> torch.Tensor:
        p = params.detach().to("cpu", torch.float64).numpy()
        e, _ = ansatz.forward(p)
        ctx.ansatz = ansatz
        ctx.save_for_backward(params)
        return torch.tensor(e, dtype=torch.float64, device=params.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (params,) = ctx.saved_tensors
        g, _ = ctx.ansatz.backward(params.detach().to("cpu", torch.float64).numpy())
        grad = torch.from_numpy(g).to(device=params.device, dtype=params.dtype)
        return grad_out.to(grad.dtype) * grad, None


def expectation(params: torch.Tensor, ansatz: Ansatz) -> torch.Tensor:
    return Expectation.apply(params, ansatz)
