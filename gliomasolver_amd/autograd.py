"""torch.autograd.Function over the C ABI's forward + adjoint, for gradcheck.

theta = (Dw_phys [cm^2/day], rho [1/day], T [day], x0, y0, z0 [cells]) as a
float64 tensor of shape (6,).  The forward runs the HIP forward solve and the
hand-written adjoint (gh_rd_adjoint) once and keeps the gradient; backward
scales it by the incoming gradient.  Torch is plumbing here: all compute is
in libgliomahip.so.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import Context, FIELD_PG, FIELD_PHI, FIELD_PW, gaussian_ic

LIK_DEFAULT = Context.LIK_DEFAULT


@dataclass
class AdjointProblem:
    """Fixed inputs of one likelihood evaluation on a uniform grid."""
    ctx: Context
    L: float            # characteristic length: Dw = Dw_phys / L^2
    nsteps: int         # frozen step count; dt = T / nsteps
    sigma: float = 1.0  # IC width (cells)
    peak: float = 0.1   # IC peak
    ckpt_every: int = 64
    lik: tuple = LIK_DEFAULT
    stepPET: int = 1
    stepTi: int = 1

    def evaluate(self, theta: np.ndarray):
        """(J, dJ/dtheta) for theta = (Dw_phys, rho, T, x0, y0, z0)."""
        Dw_phys, rho, T = float(theta[0]), float(theta[1]), float(theta[2])
        x0 = [float(v) for v in theta[3:6]]
        c = self.ctx
        pw, pg = c.get_field(FIELD_PW), c.get_field(FIELD_PG)
        c.set_field(FIELD_PHI, gaussian_ic(c.dtype, pw, pg, x0, self.sigma, self.peak))
        Dw = Dw_phys / self.L ** 2
        c.rd_prepare(Dw, 0.1 * Dw, rho)
        loss, g = c.rd_adjoint(self.nsteps, T, x0, self.sigma, self.ckpt_every, self.lik, self.stepPET, self.stepTi)
        grad = np.array([g[0] / self.L ** 2, g[1], g[2], g[3], g[4], g[5]])
        return float(loss[0]), grad


class NegLogLikelihood(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta: torch.Tensor, problem: AdjointProblem):
        J, grad = problem.evaluate(theta.detach().cpu().numpy())
        ctx.save_for_backward(torch.as_tensor(grad, dtype=theta.dtype))
        return torch.as_tensor(J, dtype=theta.dtype)

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return g * grad, None


def neg_log_likelihood(theta: torch.Tensor, problem: AdjointProblem) -> torch.Tensor:
    return NegLogLikelihood.apply(theta, problem)
