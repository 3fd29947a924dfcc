"""torch.autograd binding of the ABEA best-path score.

``abea_score(plan, scalings, model)`` returns score[r], the f5c best-path
score of read r (the float end-search maximum, GPU/align_cuda.cl:722-746),
as a differentiable function of the per-read scalings and the pore model.
The forward runs the plan's pre -> core -> post; the backward runs the
hand-written adjoint kernel (``abea_plan_adjoint``), which sums the
emission derivatives over the D and U cells of each traced path in double.

The gradient is the one of the score along the path the forward traced
(the best path is piecewise constant in the parameters).  Columns:

  scalings [n, 4] = (scale, shift, var, log_var): d/dscale, d/dshift; var
      and log_var get 0, since f5c's emission ignores them
      (GPU/align_cuda.cl:95, 515).
  model [4096, 3] = (level_mean, level_stdv, level_log_stdv): the partial
      derivatives of the function the forward computes, which reads the
      cached level_log_stdv as an input of its own (GPU/align_cuda.cl:
      517-518): d/dlevel_stdv holds the log column fixed, and the log column
      gets d/dlevel_log_stdv (-1 per D or U cell of the k-mer, weighted).  A
      caller that ties level_log_stdv = log(level_stdv) in torch gets the total
      derivative from the chain rule.
"""
from __future__ import annotations

import torch

from .engine import Plan


class ABEAScore(torch.autograd.Function):
    @staticmethod
    def forward(ctx, scalings: torch.Tensor, model: torch.Tensor, plan: Plan):
        plan.set_params(scalings, model)
        plan.forward()
        ctx.plan = plan
        ctx.generation = plan.generation
        ctx.n = plan.batch.n_reads
        ctx.dtypes = (scalings.dtype, model.dtype)
        return plan.score[:ctx.n].clone()

    @staticmethod
    def backward(ctx, grad_score: torch.Tensor):
        plan: Plan = ctx.plan
        if plan.generation != ctx.generation:
            raise RuntimeError("abea_score: the plan ran another forward since this one; "
                               "its backward needs this forward's paths")
        g_read, g_model = plan.adjoint(grad_score.detach().to(torch.float32))
        n = ctx.n
        gs = torch.zeros((n, 4), dtype=torch.float32, device=grad_score.device)
        gs[:, 0] = g_read[:, 1]  # scale
        gs[:, 1] = g_read[:, 0]  # shift
        gs[:, 2] = g_read[:, 2]  # var (0)
        return gs.to(ctx.dtypes[0]), g_model.to(ctx.dtypes[1]), None


def abea_score(plan: Plan, scalings: torch.Tensor, model: torch.Tensor) -> torch.Tensor:
    """Best-path score per read, differentiable in scalings and model."""
    return ABEAScore.apply(scalings, model, plan)
