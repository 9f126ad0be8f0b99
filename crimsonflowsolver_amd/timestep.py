"""The generalized-alpha timestep of a cylinder with a resistance outlet on the
GPU: the nonlinear loop of itrdrv (itrdrv.f90:1284-1528) over the device
entry points, one instance.

Per timestep (itrPredict with ipred = 1, itrPC.f90:104-108; itrBC):
    y = y_n, ac = ac_n (gamma - 1)/gamma; itrBC sets the inflow and walls
    for each nonlinear iteration:
        itrYAlpha (itrPC.f90:326-334); ElmGMR into the memLS matrix (interior
        and boundary elements); bc3Res; ElmpvsQ of the outlet resistance
        (elmgmr.f90:1159-1176); memLS NS solve with the outlet face coupled
        (faceRes = R alpha_f gamma Delt, elmgmr.f90:1806); itrCorrect
        (itrPC.f90:176-181); itrBC.

Inlet and wall nodes carry the three velocity codes (inlet: parabolic
profile of the mean velocity `inflow(t)`, walls: no slip); the outlet is
natural with the resistance surface on its nodes that carry no code.  With a
deformable wall (`wall`, ideformwall = 1) only the inlet nodes carry codes;
the wall nodes move with the displacement u, predicted, set at alpha_f and
corrected with the step (itrPC.f90:127-132, 183-186, 333-336).

Tensors are torch tensors on the GPU in the C ABI's column-major layouts:
y and ac as (4, nshg) C-ordered (y(nshg,4) Fortran), Res4 as (nshg, 4).
"""
from __future__ import annotations

import numpy as np

from . import (BC_TYPE_Dir, BC_TYPE_Neu, ElmGMRInterior, LS_TYPE_NS, form_params, genadj, memLS_BC_CREATE,
               memLS_LHS_CREATE, memLS_LS_CREATE, memLS_SOLVE_dev)
from . import mesh as M


def ns_solver(relTol=1e-8, dimKry=100, maxItr=10, relTolIn=(0.01, 0.2), maxItrIn=(1, 200)):
    """memLS NS solver; the defaults are the 1.3M-node workload's (Krylov
    dimension 100, tolerance 1e-8, 10 outer iterations)."""
    return memLS_LS_CREATE(LS_TYPE_NS, relTol=relTol, dimKry=dimKry, maxItr=maxItr, relTolIn=list(relTolIn),
                           maxItrIn=list(maxItrIn))


class ResistanceCylinder:
    """One cylinder (mesh.CylinderMesh) with Dirichlet inlet and walls and a
    resistance outlet, stepped on the GPU.  `params` holds rho, mu, Delt and
    the generalized-alpha coefficients (form_params).  `bct`, a bct.BCt
    attached to this mesh, replaces the parabolic inflow by the time series
    it holds (BCint)."""

    def __init__(self, mesh, R=1333.0, params=None, inflow=lambda t: 10.0, boundary=True, ls=None,
                 transpose=False, bct=None, wall=None, nodetag=None, uref=None):
        import torch
        self.torch = torch
        self.dev = torch.device("cuda:0")
        self.mesh = m = mesh
        self.N = N = m.numnp
        self.R = float(R)
        self.params = p = params if params is not None else form_params(ideformwall=1 if wall else 0)
        self.inflow = inflow
        self.bct = bct
        self.wall = wall is not None
        if self.wall and (p.ideformwall != 1 or not boundary):
            raise ValueError("a deformable wall needs ideformwall = 1 and the boundary elements")
        self.colm, self.rowp = genadj(m.ien, N)
        if self.wall:
            self.ibc = np.where(m.inlet, M.IBC_VEL, 0).astype(np.int32)
        else:
            self.ibc, _ = M.dirichlet_codes(m)
        self.form = ElmGMRInterior(m.ien, self.colm, self.rowp, N, iBC=self.ibc, params=p)
        if boundary:
            self.form.set_boundary(*M.cylinder_boundary(m))
        if self.wall:
            self.form.set_wall(wall, nodetag=nodetag, uref=uref)
            self.ua = torch.zeros((3, N), dtype=torch.float64, device=self.dev)
            self.form.set_displacement(self.ua)
        self.fixed = np.nonzero(self.ibc)[0]
        self.out = np.nonzero(m.outlet & (self.ibc == 0))[0]
        self.nabi = m.nabi()
        self.lhs = self._lhs()
        # a second matrix on the same graph and faces for the adjoint's
        # transpose solve
        self.lhsT = self._lhs() if transpose else None
        self.ls = ls if ls is not None else ns_solver()
        T = self.T
        self.dx = T(m.x)
        self.dibc = torch.from_numpy(self.ibc.astype(np.int32)).to(self.dev)
        self.dglob = torch.from_numpy(self.out.astype(np.int32)).to(self.dev)
        self.dnabi = torch.from_numpy(np.ascontiguousarray(self.nabi[self.out])).to(self.dev)  # (nf,3)
        self.dq = torch.zeros(1, dtype=torch.float64, device=self.dev)
        self.res4 = torch.zeros((N, 4), dtype=torch.float64, device=self.dev)
        rr = m.x[:, 0] ** 2 + m.x[:, 1] ** 2
        self._inl = m.inlet & ~m.wall
        self._prof = 2.0 * np.clip(1.0 - rr[self._inl] / m.radius ** 2, 0.0, None)
        # free degrees of freedom of the solve: every row but the fixed
        # nodes' velocity components, as a (N, 4) mask in the Res4 layout
        mask = np.ones((N, 4))
        mask[self.fixed, :3] = 0.0
        self.mask4 = torch.from_numpy(mask).to(self.dev)

    def _lhs(self):
        p = self.params
        lhs = memLS_LHS_CREATE(self.N, self.N, self.rowp.size, None, self.colm, self.rowp, 2)
        memLS_BC_CREATE(lhs, 1, self.fixed.size, 3, BC_TYPE_Dir, self.fixed + 1,
                        np.zeros((3, self.fixed.size), order="F"))
        memLS_BC_CREATE(lhs, 2, self.out.size, 3, BC_TYPE_Neu, self.out + 1, np.asfortranarray(self.nabi[self.out].T))
        return lhs

    def T(self, a):
        """Host (n, k) array -> device (k, n) C tensor (its Fortran layout)."""
        return self.torch.from_numpy(np.ascontiguousarray(np.asarray(a, order="F").T)).to(self.dev)

    @property
    def faceres(self):
        p = self.params
        return np.array([0.0, self.R * p.alfi * p.gami * p.Delt])

    def bc(self, t):
        """BC(nshg,7) on the device as (7, nshg): BC(:,5) the inlet profile
        of inflow(t), zero elsewhere (walls); with `bct`, BC(:,3:5) as BCint
        sets them at time t."""
        bc = np.zeros((self.N, 7), order="F")
        if self.bct is not None:
            self.bct.apply(t, bc)
        else:
            bc[self._inl, 4] = self.inflow(t) * self._prof
        return self.T(bc)

    def state(self, seed=20240611):
        """The seeded Poiseuille state (mesh.poiseuille_state) on the device."""
        y, ac = M.poiseuille_state(self.mesh, seed=seed)
        return self.T(y), self.T(ac)

    def residual(self, yold, acold, y, ac, ya=None, aca=None, lhs=None, uold=None, u=None):
        """Res4 of the current iterate (formation into `lhs`, default the
        case's matrix; bc3Res; ElmpvsQ) in self.res4; returns its norm.  With
        a deformable wall, uold and u (3, nshg) give the displacement."""
        torch = self.torch
        if ya is None:
            ya, aca = torch.empty_like(y), torch.empty_like(ac)
        f = self.form
        if self.wall:
            if uold is None or u is None:
                raise ValueError("a deformable wall's residual needs uold and u")
            f.wall_alpha(uold, u, self.ua)
        f.yalpha(yold, acold, y, ac, ya, aca)
        f.sync()
        f.to_lhs(ya, aca, self.dx, self.res4, lhs if lhs is not None else self.lhs)
        f.elmpvsq(self.dglob, self.dnabi, self.R, ya, self.res4, self.dq)  # ElmpvsQ, then bc3Res (elmgmr.f90:528-555)
        f.bc3res(self.res4, res4_layout=True)
        f.sync()
        return float(torch.linalg.vector_norm(self.res4))

    def step_wall(self, yold, acold, uold, t, **kw):
        """step() with the deformable wall's displacement: returns (y, ac,
        u, residual norms)."""
        if not self.wall:
            raise ValueError("no deformable wall")
        return self._step(yold, acold, t, uold=uold, **kw)

    def step(self, yold, acold, t, iters=3, tol=None, max_iters=200, unit_rhs=False, atol=0.0):
        """One timestep from (yold, acold) to time t.  With tol None, `iters`
        nonlinear iterations; otherwise iterations until the residual norm is
        at most max(tol times its first value, atol) (at most max_iters).
        unit_rhs
        scales each solve's right-hand side to unit norm (and the increment
        back): memLS's GE rejects pivots below 2 eps in absolute terms
        (memLS.f90:641-642), which the Gram matrix of a residual driven far
        down reaches.  Returns (y, ac, residual norms)."""
        if self.wall:
            raise ValueError("a deformable wall steps with step_wall")
        y, ac, _, norms = self._step(yold, acold, t, iters=iters, tol=tol, max_iters=max_iters, unit_rhs=unit_rhs,
                                     atol=atol)
        return y, ac, norms

    def _step(self, yold, acold, t, uold=None, iters=3, tol=None, max_iters=200, unit_rhs=False, atol=0.0):
        torch = self.torch
        f = self.form
        y, ac = torch.empty_like(yold), torch.empty_like(acold)
        ya, aca = torch.empty_like(yold), torch.empty_like(acold)
        u = None
        dBC = self.bc(t)
        f.predict(yold, acold, y, ac)
        if self.wall:
            u = uold.clone()
            f.wall_predict(uold, yold, acold, u)
        f.itr_bc(self.dibc, dBC, y)
        norms = []
        n = iters if tol is None else max_iters
        for k in range(n + (0 if tol is None else 1)):
            norms.append(self.residual(yold, acold, y, ac, ya, aca, uold=uold, u=u))
            if tol is not None and (norms[-1] <= max(tol * norms[0], atol) or k == n):
                break
            if unit_rhs:
                self.res4 /= norms[-1]
            memLS_SOLVE_dev(self.lhs, self.ls, 4, self.res4, res=self.faceres)
            if unit_rhs:
                self.res4 *= norms[-1]
            f.correct(y, ac, self.res4)
            if self.wall:
                f.wall_correct(u, self.res4)
            f.itr_bc(self.dibc, dBC, y)
            f.sync()
        return y, ac, u, norms
