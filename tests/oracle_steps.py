"""Generalized-alpha timesteps of the resistance-outlet cylinder driven by the
oracle (the CPU restatement): the same steps as
crimsonflowsolver_amd.timestep.ResistanceCylinder.step, on host arrays, for
the GPU tests that need the recurrence computed independently of the HIP
path (test infrastructure, not a product path)."""
import numpy as np


class OracleCylinder:
    """With `wall` (oracle.make_wall), the vessel wall is deformable: only
    the inlet nodes carry velocity codes and the wall nodes' displacement
    follows itrPC.f90's predictor, alpha_f level and corrector."""

    def __init__(self, O, mesh, R=1333.0, mu=0.04, inflow=lambda t: 10.0, relTol=1e-6, wall=None, nodetag=None,
                 uref=None):
        from crimsonflowsolver_amd import mesh as M
        self.O, self.m = O, mesh
        self.N = N = mesh.numnp
        self.R = float(R)
        self.p = O.make_params(mu=mu)
        self.mu = mu
        self.colm, self.rowp = O.genadj(mesh.ien, N)
        self.wall, self.tag, self.uref = wall, nodetag, uref
        if wall is not None:
            self.ibc = np.where(mesh.inlet, M.IBC_VEL, 0).astype(np.int32)
        else:
            self.ibc, _ = M.dirichlet_codes(mesh)
        self.bnd = M.cylinder_boundary(mesh)
        self.nabi = mesh.nabi()
        self.out = np.nonzero(mesh.outlet & (self.ibc == 0))[0]
        self.fixed = np.nonzero(self.ibc)[0]
        p = self.p
        self.faceres = np.array([0.0, self.R * p.alfi * p.gami * p.Delt])
        self.ls = O.MemLS(self.colm, self.rowp, 2, relTol=relTol, dimKry=100, maxItr=10,
                          relTolIn=[1e-3, 1e-3], maxItrIn=[10, 500])
        self.ls.bc_create(1, self.fixed + 1, 3, O.BC_TYPE_Dir, np.zeros((3, self.fixed.size), order="F"))
        self.ls.bc_create(2, self.out + 1, 3, O.BC_TYPE_Neu, np.asfortranarray(self.nabi[self.out].T))
        rr = mesh.x[:, 0] ** 2 + mesh.x[:, 1] ** 2
        self.inl = mesh.inlet & ~mesh.wall
        self.prof = 2.0 * np.clip(1.0 - rr[self.inl] / mesh.radius ** 2, 0.0, None)
        self.inflow = inflow

    def bc(self, t):
        bc = np.zeros((self.N, 7), order="F")
        bc[self.inl, 4] = self.inflow(t) * self.prof
        return bc

    def residual(self, yold, acold, y, ac, uold=None, u=None):
        O, p, m = self.O, self.p, self.m
        ya, aca = O.itr_yalpha(yold, acold, y, ac, p.alfi, p.almi)
        res, K, P = O.elmgmr_interior(m.ien, ya, aca, m.x, self.colm, self.rowp, p)
        if self.wall is not None:
            ua = uold + p.alfi * (u - uold)  # itrYAlpha (itrPC.f90:333-336), on every node here
            res = O.boundary_wall(*self.bnd, ya, aca, ua, m.x, res, self.wall, uref=self.uref, tag=self.tag,
                                  mu=self.mu, ibc=self.ibc, bc=np.zeros((self.N, 6)), colm=self.colm,
                                  rowp=self.rowp, lhsK=K)
        else:
            res = O.boundary_res(*self.bnd, ya, m.x, res, mu=self.mu)
        res = O.bc3res(self.ibc, res)
        res, _ = O.elmpvsq_resistance(res, ya, self.out, self.nabi[self.out], self.R)
        return res, K, P

    def step(self, yold, acold, t, tol=1e-12, max_iters=200, atol=0.0, uold=None):
        """Iterations until |res| <= max(tol |res_0|, atol); returns (y, ac,
        norms), or (y, ac, u, norms) with a deformable wall."""
        O, p = self.O, self.p
        bc = self.bc(t)
        y, ac = O.itr_predict(yold, acold, p.gami)
        u = None
        if self.wall is not None:
            betai, wn = 1.0, self.wall_nodes
            u = uold.copy()
            u[wn] = (uold[wn] + p.Delt * yold[wn, :3] +
                     0.5 * ((p.gami - 2.0 * betai) / p.gami) * p.Delt * p.Delt * acold[wn, :3])
        y = O.itr_bc(self.ibc, bc, y)
        norms = []
        for k in range(max_iters + 1):
            res, K, P = self.residual(yold, acold, y, ac, uold, u)
            norms.append(float(np.linalg.norm(res)))
            if norms[-1] <= max(tol * norms[0], atol) or k == max_iters:
                break
            # the right-hand side scaled to unit norm: memLS's GE rejects
            # pivots below 2 eps in absolute terms (memLS.f90:641-642), which
            # the Gram matrix of a small residual reaches
            sol = self.ls.solve(np.asfortranarray(res.T) / norms[-1], O.pack_val4(self.colm, self.rowp, K, P),
                                res=self.faceres) * norms[-1]
            y, ac = O.itr_correct(y, ac, sol, p.gami, p.alfi, p.Delt)
            if self.wall is not None:
                u[self.wall_nodes] += p.Delt * p.Delt * 1.0 * np.asarray(sol).T[self.wall_nodes, :3]
            y = O.itr_bc(self.ibc, bc, y)
        if self.wall is not None:
            return y, ac, u, norms
        return y, ac, norms

    @property
    def wall_nodes(self):
        """The deformable wall's nodes: those of the boundary faces with
        iBCB bit 4 (gendat.f90:172-225)."""
        ienb, iBCB, _ = self.bnd
        f = np.asarray(ienb)[(np.asarray(iBCB)[:, 0] & 16) != 0, :3] - 1
        return np.unique(f)
