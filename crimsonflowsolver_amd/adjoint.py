"""Adjoint harness over the GPU formation (torch autograd on device tensors).

``InteriorResidual`` wraps the interior residual of ``ElmGMRInterior``
(elmgmr.f90:82-255, before bc3Res) as a ``torch.autograd.Function`` of the
nodal state y(nshg,4) = (u1,u2,u3,p) and the viscosity mu; its backward is
the hand-written VJP kernel (``crimson_form_vjp_dev``).  The viscosity of a
formation handle is fixed at setup, so ``mu`` is passed as a 0-d tensor that
must equal the handle's value: its gradient is the VJP's mubar.

``resistance_term`` is the ElmpvsQ contribution of a resistance outlet
(elmgmr.f90:1159-1176): res_i(1:3) += -R Q NABI_i with Q = sum_j NABI_j . u_j,
differentiable in u and R by torch itself (small: face nodes only).

``checkpointed_adjoint`` runs a forward recurrence y_{n+1} = step(y_n, theta)
for nsteps keeping every k-th state, and accumulates the reverse sweep by
recomputing each segment from its checkpoint (memory O(nsteps/k + k) states
instead of O(nsteps)).

Tensors are torch tensors on the GPU in the column-major layouts of the C ABI
stored as (4, nshg) C-ordered views transposed, i.e. ``y_f = y_dev.T`` is the
Fortran (nshg,4) array; helpers below keep that convention.
"""
from __future__ import annotations

import torch


def _fortran(t):
    """(n, k) tensor -> contiguous storage of its Fortran (column-major) layout
    as a (k, n) C tensor."""
    return t.t().contiguous()


class InteriorResidual(torch.autograd.Function):
    """res = R(y; mu) with y, res (nshg,4) on the GPU."""

    @staticmethod
    def forward(ctx, y, mu, form, ac, x):
        if abs(float(mu) - form.params.mu) > 1e-15 * abs(form.params.mu):
            raise ValueError("mu must equal the formation handle's viscosity")
        yf, acf, xf = _fortran(y), _fortran(ac), _fortran(x)
        resf = torch.empty_like(yf)
        form.dev(yf, acf, xf, resf, lhs=0)
        ctx.form = form
        ctx.save_for_backward(yf, acf, xf)
        return resf.t().contiguous()

    @staticmethod
    def backward(ctx, w):
        yf, acf, xf = ctx.saved_tensors
        wf = _fortran(w.to(torch.float64))
        ybf = torch.empty_like(yf)
        mubar = ctx.form.vjp(yf, acf, xf, wf, ybf)
        return ybf.t().contiguous(), torch.tensor(mubar, dtype=torch.float64, device=w.device), None, None, None


def interior_residual(form, y, mu, ac, x):
    return InteriorResidual.apply(y, mu, form, ac, x)


def resistance_term(y, R, glob, nabi):
    """ElmpvsQ residual of one resistance face on the face nodes ``glob``
    (0-based long tensor) with NABI (nface,3): returns the (nshg,4) addition."""
    u = y[glob, :3]
    Q = (nabi * u).sum()
    add = torch.zeros_like(y)
    add = add.index_add(0, glob, torch.nn.functional.pad(-R * Q * nabi, (0, 1)))
    return add


def checkpointed_adjoint(step, y0, theta, nsteps, every, loss):
    """Gradient of loss(y_N) w.r.t. y0 and theta for y_{n+1} = step(y_n, theta).

    Keeps y at steps 0, every, 2 every, ...; the reverse sweep recomputes
    each segment from its checkpoint under autograd and back-propagates the
    running cotangent through it.  Returns (loss value, dL/dy0, dL/dtheta)."""
    ckpt = {}
    y = y0.detach()
    with torch.no_grad():
        for n in range(nsteps):
            if n % every == 0:
                ckpt[n] = y.clone()
            y = step(y, theta)
    yN = y.detach().requires_grad_(True)
    L = loss(yN)
    (ybar,) = torch.autograd.grad(L, yN)
    tbar = torch.zeros_like(theta)
    for s0 in sorted(ckpt, reverse=True):
        s1 = min(s0 + every, nsteps)
        yseg = ckpt[s0].detach().requires_grad_(True)
        th = theta.detach().requires_grad_(True)
        z = yseg
        for _ in range(s0, s1):
            z = step(z, th)
        gy, gt = torch.autograd.grad(z, (yseg, th), grad_outputs=ybar, allow_unused=True)
        ybar = gy
        if gt is not None:
            tbar = tbar + gt
    return float(L), ybar, tbar
