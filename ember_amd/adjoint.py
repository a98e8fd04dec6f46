"""Differentiable chemistry production stage.

``QssStage`` is a ``torch.autograd.Function`` over the library's discrete adjoint of the
batched QSS stage (``emx_source_adjoint_forward`` / ``emx_source_adjoint_backward``).  The
forward pass integrates every point over ``dt`` exactly as the production stage does
(SourceSystemQSS::integrateToTime, src/sourceSystem.cpp:434-443) and returns the new
states and the heat release integrated over the accepted internal steps.  The backward
pass returns gradients with respect to the input states and to the natural logs of the
reaction rate multipliers (Cantera m_perturb, CanteraGas::setRateMultiplier,
src/chemistry0d.cpp:634-642), with the internal step sequence held fixed.  The library
checkpoints the integrator every 16 accepted steps in the forward pass and re-integrates
between checkpoints in the backward pass.
"""
from __future__ import annotations

import numpy as np
import torch

from . import capi


class _multipliers:
    """Sets the context's per-reaction multipliers to exp(ln_mult) for the adjoint calls
    and restores the caller's multipliers afterwards (emx_set_rate_multipliers waits for
    the work queued on the context stream before it replaces them), so that later
    production stages on the same context are unaffected."""

    def __init__(self, emx, ln_mult):
        self.emx = emx
        self.mult = np.exp(ln_mult.detach().double().cpu().numpy())

    def __enter__(self):
        self.prev = self.emx.rate_multipliers
        self.emx.set_rate_multipliers(self.mult)

    def __exit__(self, *exc):
        self.emx.set_rate_multipliers(self.prev)
        return False


class QssStage(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state, ln_mult, dt, emx, split_const=None, qloss=None):
        if not state.is_cuda:
            raise capi.EmberAmdError("QssStage needs the states on the GPU")
        state = state.detach().contiguous()
        n = state.shape[0]
        out = torch.empty_like(state)
        Q = torch.empty(n, dtype=torch.float64, device=state.device)
        status = torch.empty(n, dtype=torch.int32, device=state.device)
        with _multipliers(emx, ln_mult):
            emx.adjoint_forward(state, out, dt, heat_release=Q, status=status,
                                split_const=split_const, qloss=qloss)
        emx._adjoint_generation = getattr(emx, "_adjoint_generation", 0) + 1
        ctx.generation = emx._adjoint_generation
        ctx.emx, ctx.dt, ctx.split_const, ctx.qloss = emx, dt, split_const, qloss
        ctx.save_for_backward(state, ln_mult)
        ctx.mark_non_differentiable(status)
        return out, Q, status

    @staticmethod
    def backward(ctx, g_out, g_Q, g_status):
        state, ln_mult = ctx.saved_tensors
        emx = ctx.emx
        n = state.shape[0]
        Qbar = (g_Q if g_Q is not None else torch.zeros(n, dtype=torch.float64, device=state.device))
        Qbar = Qbar.to(torch.float64).contiguous()
        ybar = None if g_out is None else g_out.to(torch.float64).contiguous()
        grad_state = torch.empty_like(state)
        need_k = ctx.needs_input_grad[1]
        grad_k = (torch.empty((n, emx.n_reactions), dtype=torch.float64, device=state.device)
                  if need_k else None)
        with _multipliers(emx, ln_mult):
            if getattr(emx, "_adjoint_generation", 0) != ctx.generation:
                # another forward pass replaced the checkpoints: recompute them for this one
                out = torch.empty_like(state)
                emx.adjoint_forward(state, out, ctx.dt, split_const=ctx.split_const, qloss=ctx.qloss)
                emx._adjoint_generation += 1
                ctx.generation = emx._adjoint_generation
            emx.adjoint_backward(state, ctx.dt, grad_state, grad_lnk=grad_k, Qbar=Qbar, ybar=ybar,
                                 split_const=ctx.split_const, qloss=ctx.qloss)
        gk = None
        if need_k:
            gk = grad_k.sum(dim=0).to(device=ln_mult.device, dtype=ln_mult.dtype)
        return grad_state, gk, None, None, None, None


def qss_stage(state, ln_mult, dt, emx, split_const=None, qloss=None):
    """Differentiable production stage: returns (state_out, heat_release, status)."""
    return QssStage.apply(state, ln_mult, dt, emx, split_const, qloss)


def per_point_gradients(emx, state, dt, Qbar=None, ybar=None, split_const=None, qloss=None):
    """Per-point gradients of J_j = Qbar_j Q_j + ybar_j . y_j: (state_out, Q, grad_state,
    grad_lnk (nPoints x nReactions), status), all on the GPU."""
    n = state.shape[0]
    dev = state.device
    out = torch.empty_like(state)
    Q = torch.empty(n, dtype=torch.float64, device=dev)
    st = torch.empty(n, dtype=torch.int32, device=dev)
    emx.adjoint_forward(state, out, dt, heat_release=Q, status=st, split_const=split_const,
                        qloss=qloss)
    emx._adjoint_generation = getattr(emx, "_adjoint_generation", 0) + 1
    gs = torch.empty_like(state)
    gk = torch.empty((n, emx.n_reactions), dtype=torch.float64, device=dev)
    st2 = torch.empty(n, dtype=torch.int32, device=dev)
    emx.adjoint_backward(state, dt, gs, grad_lnk=gk, Qbar=Qbar, ybar=ybar, status=st2,
                         split_const=split_const, qloss=qloss)
    return out, Q, gs, gk, torch.minimum(st, st2)
