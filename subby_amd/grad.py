"""Hand-written adjoint behind torch.autograd.Function (subby/jax/vjp.py:35-136).

The forward pass is the FP32 inside pass; the backward pass runs the fused
inside/outside/adjoint kernel once: per branch n it reduces
sum_c g_c scale(n,c) sum_k D~_k mu_k e^{mu_k t_n} U~_k over columns, and for a
rate parameter theta it contracts the unmasked counts with the Fisher weights
(Holmes & Rubin 2002, eqs. 3-4).  Inside vectors are recomputed per wave tile
instead of being stored for every column.
"""
from __future__ import annotations

import numpy as np
import torch

from . import api
from .models import Tree, ensure_diag, hky85_diag, hky85_log_kappa_weights


class _LogLikeDist(torch.autograd.Function):
    @staticmethod
    def forward(ctx, distances, alignment, parent, model, precision, engine):
        tree = Tree(parent, distances.detach().cpu().numpy())
        ll = api.LogLike(alignment, tree, model, precision=precision, engine=engine)
        ctx.save_for_backward(distances)
        ctx.args = (alignment, parent, model, precision, engine)
        return ll

    @staticmethod
    def backward(ctx, g):
        (distances,) = ctx.saved_tensors
        alignment, parent, model, precision, engine = ctx.args
        tree = Tree(parent, distances.detach().cpu().numpy())
        dgrad, _, _ = api.loglike_grad(alignment, tree, model, g=g.detach().to(torch.float64),
                                       precision=precision, engine=engine)
        return dgrad.to(distances.device, distances.dtype), None, None, None, None, None


def LogLikeCustomGrad(alignment, tree, model, maxChunkSize: int = 128, method: str = 'eigen',
                      precision: str = 'fp32', engine=None):
    """LogLike differentiable in tree.distanceToParent (subby/jax/__init__.py:294-335).

    ``method='pade'`` (the reference's expm forward / Q M(t) backward) gives the
    same values and is served by the same eigenbasis adjoint kernels: the model
    must be diagonalizable (every DiagModel, IrrevDiagModel or RateModel this
    module builds is)."""
    if method not in ('eigen', 'pade'):
        raise ValueError("method must be 'eigen' or 'pade'")
    d = tree.distanceToParent
    if not isinstance(d, torch.Tensor):
        d = torch.as_tensor(np.asarray(d, dtype=np.float64))
    parent = np.asarray(tree.parentIndex.detach().cpu().numpy() if isinstance(tree.parentIndex, torch.Tensor)
                        else tree.parentIndex, dtype=np.int32)
    return _LogLikeDist.apply(d, alignment, parent, ensure_diag(model), precision, engine)


class _HKY85LogLike(torch.autograd.Function):
    @staticmethod
    def forward(ctx, distances, log_kappa, alignment, parent, pi, precision, engine):
        kappa = float(torch.exp(log_kappa.detach()).cpu())
        model = hky85_diag(kappa, pi)
        tree = Tree(parent, distances.detach().cpu().numpy())
        ll = api.LogLike(alignment, tree, model, precision=precision, engine=engine)
        ctx.save_for_backward(distances, log_kappa)
        ctx.args = (alignment, parent, pi, precision, engine, kappa)
        return ll

    @staticmethod
    def backward(ctx, g):
        distances, log_kappa = ctx.saved_tensors
        alignment, parent, pi, precision, engine, kappa = ctx.args
        model = hky85_diag(kappa, pi)
        tree = Tree(parent, distances.detach().cpu().numpy())
        W = hky85_log_kappa_weights(kappa, pi)
        dgrad, rgrad, _ = api.loglike_grad(alignment, tree, model, g=g.detach().to(torch.float64),
                                           rate_weights=W[None], precision=precision, engine=engine)
        return (dgrad.to(distances.device, distances.dtype),
                rgrad[0].to(log_kappa.device, log_kappa.dtype).reshape(log_kappa.shape),
                None, None, None, None, None)


def hky85_loglike(alignment, tree, log_kappa, pi, precision: str = 'fp32', engine=None):
    """Per-column HKY85 log-likelihood differentiable in the branch lengths
    (tree.distanceToParent as a tensor) and in log(kappa)."""
    d = tree.distanceToParent
    if not isinstance(d, torch.Tensor):
        d = torch.as_tensor(np.asarray(d, dtype=np.float64))
    if not isinstance(log_kappa, torch.Tensor):
        log_kappa = torch.tensor(float(log_kappa), dtype=torch.float64)
    parent = np.asarray(tree.parentIndex, dtype=np.int32)
    return _HKY85LogLike.apply(d, log_kappa, alignment, parent, np.asarray(pi, dtype=np.float64),
                               precision, engine)
