"""torch.autograd.Function around the device solve and its adjoint (used for gradient checks).

``HeldoutLoss.apply(log_eps, theta, problem, mask, params)`` solves the gamma QP on the device
for epsilon = exp(log_eps) and the given theta, and returns the held-out reconstruction loss
J = sum_{t in mask} (x_t - sum_k theta_k gamma*_kt)^2; backward() returns the adjoint gradients
from pasc_adjoint.  Tensors live on the CPU in float64; the computation runs in libpasc_hip.so.
"""
from __future__ import annotations

import numpy as np
import torch


class HeldoutLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, log_eps, theta, problem, mask, params):
        problem.set_epssqr(float(torch.exp(log_eps).item()))
        problem.set_theta(theta.detach().cpu().numpy().astype(np.float64))
        problem.set_gamma_uniform()
        problem.gamma_solve(**params)
        loss, dle, dth, _ = problem.adjoint(mask)
        ctx.save_for_backward(torch.tensor(dle, dtype=torch.float64),
                              torch.tensor(dth, dtype=torch.float64))
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        dle, dth = ctx.saved_tensors
        return grad_out * dle, grad_out * dth, None, None, None
