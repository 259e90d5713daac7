"""torch.autograd binding of the solve and its hand-written adjoint.

`refine(s, carrier, chi0, log_lambda_y, log_lambda_c, engine, ...)` runs the
GPU IRLS + CG solve (VariationalRefine, 2-D) on a batch of planes and returns
(Y, chi); its backward is the GPU adjoint (hvd_gpu_refine_adjoint: one
adjoint CG per IRLS stage with the forward's operator, stage weights
recomputed from the saved stage-start chi), giving gradients with respect to
S, chi0 and the two regularisation weights (in log space).  The carrier is
an input without gradient.
"""
from __future__ import annotations

import math

import torch

from . import gpu


def _planes(s, carrier, chi, y, sample_rate):
    return [gpu.Engine.plane(s[i], chi[i], y=None if y is None else y[i], carrier=carrier[i],
                             sample_rate=sample_rate) for i in range(s.shape[0])]


class RefineFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, s, carrier, chi0, log_lambda_y, log_lambda_c, engine, solver_kw, adj_kw):
        s = s.detach().contiguous()
        carrier = carrier.detach().contiguous()
        chi = chi0.detach().contiguous().clone()
        y = torch.empty_like(s)
        params = gpu.SolverParams.make(lambda_y=math.exp(float(log_lambda_y)),
                                       lambda_c=math.exp(float(log_lambda_c)), **solver_kw)
        stages = engine.variational_refine_stages(_planes(s, carrier, chi, y, 4.0 * 315.0e6 / 88.0), params)
        ctx.engine, ctx.params, ctx.adj_kw = engine, params, adj_kw
        ctx.save_for_backward(s, carrier, chi, torch.stack(stages))
        return y, chi

    @staticmethod
    def backward(ctx, gy, gchi):
        s, carrier, chi, stages = ctx.saved_tensors
        n = s.shape[0]
        gy = torch.zeros_like(s) if gy is None else gy.contiguous()
        gchi = torch.zeros_like(chi) if gchi is None else gchi.contiguous()
        planes = _planes(s, carrier, chi, None, 4.0 * 315.0e6 / 88.0)
        gs, gchi0, glam = ctx.engine.refine_adjoint(planes, ctx.params, [stages[i] for i in range(n)],
                                                    [gchi[i] for i in range(n)], [gy[i] for i in range(n)],
                                                    **ctx.adj_kw)
        # d/dlog(lambda) = lambda d/dlambda, summed over the batch
        dly = torch.tensor(float(ctx.params.lambda_y) * glam[:, 0].sum(), dtype=torch.float32, device=s.device)
        dlc = torch.tensor(float(ctx.params.lambda_c) * glam[:, 1].sum(), dtype=torch.float32, device=s.device)
        return torch.stack(gs), None, torch.stack(gchi0), dly, dlc, None, None, None


def refine(s, carrier, chi0, log_lambda_y, log_lambda_c, engine, *, irls_outer=4, cg_iterations=100, fast=0,
           cg_tol=1e-12, chroma_aniso=0.7, adj_iterations=100, adj_tol=1e-6):
    """Differentiable solve of a batch: s (n, h, w), carrier and chi0 (n, h,
    w, 2) float32 device tensors; log_lambda_y / log_lambda_c 0-d tensors.
    chroma_aniso must be forced (> 0): the AUTO value is not differentiated."""
    if not chroma_aniso > 0:
        raise ValueError("refine() needs a forced chroma_aniso")
    solver_kw = dict(irls_outer=irls_outer, cg_iterations=cg_iterations, fast=fast, cg_tol=cg_tol,
                     chroma_aniso=chroma_aniso)
    adj_kw = dict(adj_iterations=adj_iterations, adj_tol=adj_tol)
    return RefineFunction.apply(s, carrier, chi0, log_lambda_y, log_lambda_c, engine, solver_kw, adj_kw)
