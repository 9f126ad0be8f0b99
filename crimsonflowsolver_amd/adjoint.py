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


class StepAdjoint:
    """Discrete adjoint of converged generalized-alpha timesteps of a
    timestep.ResistanceCylinder (created with transpose=True).

    One step (itrPC.f90:104-108, 164-192, 295-365) from x_n = (y_n, ac_n)
    solves F(z) = 0 for z = (da, dp): with c = (gamma - 1)/gamma,
        ac_{n+1} = c ac_n + da,  u_{n+1} = u_n + gamma Delt da (u set by itrBC
        at fixed nodes, where da = 0),  p_{n+1} = p_n + alpha_f gamma Delt dp,
        y_alpha = (u_n + alpha_f (u_{n+1} - u_n), p_{n+1}),
        ac_alpha = ac_n + alpha_m (ac_{n+1} - ac_n),
    F = Res4(y_alpha, ac_alpha): the formation's residual with bc3Res and
    the outlet's ElmpvsQ term, on the free rows (the fixed nodes' velocity
    rows removed).  For a cotangent x_bar_{n+1} the step's adjoint is
        z_bar = (gamma Delt u_bar_{n+1} + ac_bar_{n+1}, alpha_f gamma Delt p_bar_{n+1}),
        F_z^T lam = z_bar,
        F_z^T lam = (alpha_f gamma Delt yb_u + alpha_m ab_u, alpha_f gamma Delt yb_p)
    with (yb, ab) = lam^T dF/d(y_alpha, ac_alpha) from the VJP kernels
    (crimson_form_vjp_ac_dev, crimson_elmpvsq_vjp_dev), and
        u_bar_n = u_bar_{n+1} - yb_u (free nodes), -(1 - alpha_f) yb_u (fixed),
        p_bar_n = p_bar_{n+1} - yb_p,
        ac_bar_n = c ac_bar_{n+1} - (1 - alpha_m + alpha_m c) ab,
        mu_bar -= lam^T dF/dmu,  R_bar -= lam^T dF/dR.
    The transpose solve is flexible GMRES on F_z^T (modified Gram-Schmidt,
    restarted), right-preconditioned by memLS's NS solver on the block
    transpose of the step's assembled tangent (crimson_lhs_transpose_dev):
    the lazy tangent K approximates -F_z, so the preconditioner is -K^-T.

    gradient() runs the forward recurrence keeping every `every`-th state,
    then the reverse sweep, recomputing each segment from its checkpoint."""

    def __init__(self, case, pre_ls=None, tol=1e-10, restart=30, max_restarts=6):
        import torch
        from .timestep import ns_solver
        if case.lhsT is None:
            raise ValueError("the case needs transpose=True (a second matrix for the transpose solve)")
        self.case, self.torch = case, torch
        self.pre_ls = pre_ls if pre_ls is not None else ns_solver(relTol=1e-3, maxItr=4)
        self.tol, self.restart, self.max_restarts = tol, restart, max_restarts
        self.cbar = torch.zeros(2, dtype=torch.float64, device=case.dev)
        self.stats = []

    # F_z^T lam, lam in the Res4 layout (N, 4); also the raw cotangents
    def _op(self, ya, aca, lam4):
        torch, c = self.torch, self.case
        p, f = c.params, c.form
        w = lam4.t().contiguous()
        yb, ab = torch.empty_like(w), torch.empty_like(w)
        mub = f.vjp_ac(ya, aca, c.dx, w, yb, ab)
        f.elmpvsq_vjp(c.dglob, c.dnabi, c.R, ya, lam4, yb, self.cbar)
        f.sync()
        k = p.alfi * p.gami * p.Delt
        out = torch.empty_like(lam4)
        out[:, :3] = (k * yb[:3] + p.almi * ab[:3]).t()
        out[:, 3] = k * yb[3]
        out *= c.mask4
        return out, yb, ab, mub

    def _pre(self, r4):
        from . import memLS_SOLVE_dev
        c = self.case
        z = r4.clone()
        memLS_SOLVE_dev(c.lhsT, self.pre_ls, 4, z, res=c.faceres)
        z *= -1.0
        return z * c.mask4

    def _solve(self, ya, aca, b4):
        """FGMRES for F_z^T lam = b4; returns (lam, iterations, relative
        residual)."""
        import numpy as np
        torch = self.torch
        bn = float(torch.linalg.vector_norm(b4))
        x = torch.zeros_like(b4)
        if bn == 0.0:
            return x, 0, 0.0
        r, beta, its = b4.clone(), bn, 0
        for _ in range(self.max_restarts):
            V, Z = [r / beta], []
            m = self.restart
            H = np.zeros((m + 1, m))
            yk = np.zeros(0)
            for j in range(m):
                z = self._pre(V[j])
                Z.append(z)
                w = self._op(ya, aca, z)[0]
                for i in range(j + 1):
                    h = float((w * V[i]).sum())
                    H[i, j] = h
                    w -= h * V[i]
                hn = float(torch.linalg.vector_norm(w))
                H[j + 1, j] = hn
                its += 1
                e1 = np.zeros(j + 2)
                e1[0] = beta
                yk = np.linalg.lstsq(H[:j + 2, :j + 1], e1, rcond=None)[0]
                est = float(np.linalg.norm(H[:j + 2, :j + 1] @ yk - e1))
                if hn == 0.0 or est <= self.tol * bn:
                    break
                V.append(w / hn)
            for i in range(yk.size):
                x += float(yk[i]) * Z[i]
            r = b4 - self._op(ya, aca, x)[0]
            beta = float(torch.linalg.vector_norm(r))
            if beta <= self.tol * bn:
                break
        return x, its, beta / bn

    def reverse_step(self, yn, acn, y1, ac1, ybar1, acbar1):
        """Cotangents of (y_n, ac_n) and (mu, R) of one converged step from
        those of (y_{n+1}, ac_{n+1}) (device tensors in the y layout)."""
        torch, c = self.torch, self.case
        p, f = c.params, c.form
        ya, aca = torch.empty_like(yn), torch.empty_like(acn)
        f.yalpha(yn, acn, y1, ac1, ya, aca)
        f.sync()
        f.to_lhs(ya, aca, c.dx, c.res4, c.lhs)
        f.sync()
        c.lhsT.transpose_from(c.lhs)
        gdt = p.gami * p.Delt
        zb = torch.empty((c.N, 4), dtype=torch.float64, device=c.dev)
        zb[:, :3] = (gdt * ybar1[:3] + acbar1[:3]).t()
        zb[:, 3] = p.alfi * gdt * ybar1[3]
        zb *= c.mask4
        lam, its, rel = self._solve(ya, aca, zb)
        self.stats.append((its, rel))
        _, yb, ab, mub = self._op(ya, aca, lam)
        rbar = float(self.cbar[0])
        free = c.mask4[:, 0]
        cf = (p.gami - 1.0) / p.gami
        ybar0 = torch.empty_like(ybar1)
        ybar0[:3] = ybar1[:3] * free - yb[:3] * (free + (1.0 - p.alfi) * (1.0 - free))
        ybar0[3] = ybar1[3] - yb[3]
        acbar0 = torch.zeros_like(acbar1)
        acbar0[:3] = cf * acbar1[:3] - (1.0 - p.almi + p.almi * cf) * ab[:3]
        return ybar0, acbar0, -mub, -rbar

    def gradient(self, y0, ac0, t0, nsteps, every, cy, ca, step_kw=None, progress=None):
        """J = <cy, y_N> + <ca, ac_N> after nsteps steps from (y0, ac0) at
        time t0 (step_kw: ResistanceCylinder.step's keywords).  Returns (J,
        dJ/dy0, dJ/dac0, dJ/dmu, dJ/dR).  progress(phase, step), when given,
        is called after the forward pass ("forward", nsteps) and after each
        checkpoint segment's reverse sweep ("reverse", its first step)."""
        c = self.case
        kw = dict(step_kw or {})
        dt = c.params.Delt
        ck = {}
        y, ac = y0, ac0
        for n in range(nsteps):
            if n % every == 0:
                ck[n] = (y.clone(), ac.clone())
            y, ac, _ = c.step(y, ac, t0 + (n + 1) * dt, **kw)
        J = float((cy * y).sum() + (ca * ac).sum())
        if progress:
            progress("forward", nsteps)
        yb, ab = cy.clone(), ca.clone()
        mub = rb = 0.0
        self.stats = []
        for s0 in sorted(ck, reverse=True):
            s1 = min(s0 + every, nsteps)
            states = [ck[s0]]
            for n in range(s0, s1):
                yy, aa, _ = c.step(*states[-1], t0 + (n + 1) * dt, **kw)
                states.append((yy, aa))
            for n in range(s1 - 1, s0 - 1, -1):
                yb, ab, dm, dr = self.reverse_step(*states[n - s0], *states[n - s0 + 1], yb, ab)
                mub += dm
                rb += dr
            del states
            if progress:
                progress("reverse", s0)
        return J, yb, ab, mub, rb
