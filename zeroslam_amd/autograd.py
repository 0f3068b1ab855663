"""torch.autograd wrapper of the converged bundle adjustment and its adjoint.

`BundleAdjust.apply(uv, K, problem)` solves the bundle adjustment of `problem`
with the observations `uv` (n_obs x 2) and intrinsics `K` (fx, fy, cx, cy) to
convergence on the GPU and returns the converged landmarks (n_points x 3);
its backward pass is the engine's adjoint (implicit differentiation through the
converged normal equations), so gradients of any loss of the landmarks flow to
`uv` and `K`.  Tensors are float64 on the CPU; torch is imported only here.
"""
from __future__ import annotations

import numpy as np
import torch

from .engine import FP64_FAST, Graph
from .problem import Problem


def solve_problem(problem: Problem, uv: np.ndarray, K: np.ndarray, precision: int = FP64_FAST) -> Graph:
    """Solve to the stationary point the adjoint differentiates: LM to
    convergence, then points without a local optimum (landmarks seen without
    parallax: their exact Hessian block is not positive definite) are held fixed
    and undamped Gauss-Newton steps polish the rest.  g.held flags the held
    points by id; the adjoint treats them as constants."""
    pr = problem.copy()
    pr.obs_uv = np.ascontiguousarray(uv, dtype=np.float64).reshape(-1, 2)
    pr.obs_K = np.tile(np.asarray(K, dtype=np.float64), (pr.n_obs, 1))
    g = Graph.from_problem(pr)
    g.solve_to_convergence(precision, polish=0)
    g.held = np.zeros(pr.n_points, np.uint8)
    for _ in range(3):
        held = g.hold_indefinite_points()
        g.held |= held
        g.gauss_newton(4)
        if not held.any():
            break
    return g


class BundleAdjust(torch.autograd.Function):
    @staticmethod
    def forward(ctx, uv: torch.Tensor, K: torch.Tensor, problem: Problem):
        g = solve_problem(problem, uv.detach().cpu().numpy(), K.detach().cpu().numpy())
        ctx.graph = g
        return torch.from_numpy(g.points())

    @staticmethod
    def backward(ctx, grad_points: torch.Tensor):
        d_obs, d_K = ctx.graph.adjoint(grad_points.detach().cpu().numpy().astype(np.float64))
        return torch.from_numpy(d_obs), torch.from_numpy(d_K), None
