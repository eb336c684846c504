"""Differentiable MHM_RP rollouts on a periodic uniform level (FP64).

Each step of the rollout is ghost-zone preparation -> one MHM_RP patch-group update -> scatter back, the
pipeline of ``UniformLevel.step`` with a fixed dt (src/Fluid/Flu_AdvanceDt.cpp:38-100 for a single level).  The three
pieces are ``torch.autograd.Function``s whose backward passes call the hand-written device kernels:

* ``PrepareFn``   backward = ``hydro_prepare_uniform_adjoint_async`` (sum of every ghost copy of a cell),
* ``MHMStepFn``   forward = the production solver kernel, backward = ``hydro_mhm_adjoint_async``, which
                  recomputes the step's internals from the saved patch-group input (per-step checkpointing:
                  only the 22^3 inputs of every step are kept),
* ``CloseFn``     backward = ghost-free preparation (the scatter's transpose is a gather).

``rollout_tangent`` pushes a tangent through the same rollout with ``hydro_mhm_tangent_async`` (forward-mode),
for dot-product tests of the adjoint.
"""
from __future__ import annotations

import torch

from . import hydro

GHOST = 3


class PrepareFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, level, ghost):
        ctx.shape, ctx.ghost = tuple(level.shape), ghost
        return hydro.prepare_uniform(level, ghost)

    @staticmethod
    def backward(ctx, g):
        return hydro.prepare_uniform_adjoint(g.contiguous(), ctx.shape, ctx.ghost), None


class CloseFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pg_out, level_shape):
        level = torch.empty(level_shape, dtype=pg_out.dtype, device=pg_out.device)
        return hydro.close_uniform(pg_out, level)

    @staticmethod
    def backward(ctx, g):
        return hydro.prepare_uniform(g.contiguous(), 0), None


class MHMStepFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pg_in, dt, dh, kw):
        out, _ = hydro.fluid_solver(pg_in, dt, dh, store_flux=False, **kw)
        ctx.save_for_backward(pg_in)
        ctx.dt, ctx.dh, ctx.kw = dt, dh, kw
        return out

    @staticmethod
    def backward(ctx, g):
        (pg_in,) = ctx.saved_tensors
        return hydro.mhm_adjoint(pg_in, g.contiguous(), ctx.dt, ctx.dh, **ctx.kw), None, None, None


def rollout(level0, nsteps, dt, dh, **kw):
    """``nsteps`` MHM_RP steps of a periodic (5, Z, Y, X) FP64 level with constant dt; differentiable in level0."""
    u = level0
    for _ in range(nsteps):
        u = CloseFn.apply(MHMStepFn.apply(PrepareFn.apply(u, GHOST), dt, dh, kw), tuple(level0.shape))
    return u


def rollout_tangent(level0, v0, nsteps, dt, dh, **kw):
    """(final level, its directional derivative along v0) by forward-mode through every step."""
    u, v = level0.contiguous(), v0.contiguous()
    ws = None
    for _ in range(nsteps):
        pu, pv = hydro.prepare_uniform(u, GHOST), hydro.prepare_uniform(v, GHOST)
        if ws is None:
            ws = hydro.adjoint_workspace(pu.shape[0], pu.device)
        out, jv = hydro.mhm_tangent(pu, pv, dt, dh, workspace=ws, **kw)
        u = hydro.close_uniform(out, torch.empty_like(u))
        v = hydro.close_uniform(jv, torch.empty_like(v))
    return u, v


def density_loss(level, target):
    """L = sum over cells of (rho - rho_target)^2."""
    return ((level[0] - target) ** 2).sum()
