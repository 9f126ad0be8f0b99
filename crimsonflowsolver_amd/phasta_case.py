"""itrdrv's time loop for a PHASTA case on the GPU: the deformable-wall
vessel with a bct.dat inflow and RCR-coupled outlets, as the reference runs
testbin/mainTests/basic (itrdrv.f90:1100-1600, solfar.f90:100-400,
elmgmr.f90:500-560, 1265-1345, 1775-1856).

Per step n (to t_{n+1} = (n+1) Delt):

* BCint at t_{n+1} fills BC's velocities (bct.BCt).
* Each RCR surface's implicit coefficients (RCR.cxx:44-88,
  AbstractBoundaryCondition.cxx:26-94): with the surface's pressure P_n and
  flow Q_n, c1 = (Rd + Rp (1 + C Rd/D)) / (1 + C Rd/D) and
  c2 = (C Rd/D (P_n + Pd(t1) - Pd(t_n) - Rp Q_n) + Pd(t1)) / (1 + C Rd/D),
  for the solve with D = alfi Delt, t1 = t_n + alfi Delt and for the update
  with D = Delt, t1 = t_{n+1}; after the step P_{n+1} = c1u Q_{n+1} + c2u.
* itrPredict (y, ac and the wall displacement), itrBC.
* Nonlinear iterations (at most `nonlinear`): itrYAlpha; the formation into
  the memLS matrix (interior, boundary elements with the deformable wall);
  ElmpvsQ (each RCR surface: Res4(1:3,i) -= (c1s Q + c2s) NABI(i), Q =
  sum NABI . u at y_alpha); bc3Res; memLS NS solve with each surface's face
  coupled by c1s alfi gami Delt; itrCorrect and the wall displacement;
  itrBC; "Residual Control": stop once sqrt(sum Res^2 / nshg) is below
  `res_criteria` after at least `min_iters` iterations (rstatic.f90).

The formation, the boundary and wall rows, the solve, ElmpvsQ and the step
updates run on the device through the C ABI; the coefficients, BCint and
the surfaces' flow and mean pressure at the end of each step (a few
reductions) are host or torch work.  NABI and NASC are the boundary
integrals of AsBNABI / AsBNASC (asbnabi.f90): sums over the faces with code
bit 1 or on a Dirichlet flow surface of N_a n WdetJb and N_a WdetJb; a
surface's nodes are those whose "global node surface number" it is.
"""
from __future__ import annotations

import numpy as np

from . import (BC_TYPE_Dir, BC_TYPE_Neu, ElmGMRInterior, LS_TYPE_NS, form_params, genadj, memLS_BC_CREATE,
               memLS_LHS_CREATE, memLS_LS_CREATE, memLS_SOLVE_dev, wall_params)
from . import bct as _bct

_TRI_A = 0.3333333333333330000000000000000000  # symtri.c:20
_TRI_B = 0.666666666666667                     # symtri.c:21
_TRI_C = 0.166666666666667                     # symtri.c:22


def surface_integrals(x, ienb, iBCB, nqb=3, dirichlet_srfs=()):
    """(NABI(nshg,3), NASC(nshg)): AsBNABI and AsBNASC over the boundary
    faces with iBCB bit 1 or whose srfID is in `dirichlet_srfs` ("List of
    Dirichlet Surfaces"), at the boundary quadrature points."""
    x = np.asarray(x, dtype=np.float64)
    ienb, iBCB = np.asarray(ienb), np.asarray(iBCB)
    N = x.shape[0]
    nabi, nasc = np.zeros((N, 3)), np.zeros(N)
    if nqb == 3:
        pts, qw = [(_TRI_B, _TRI_C, _TRI_C), (_TRI_C, _TRI_B, _TRI_C), (_TRI_C, _TRI_C, _TRI_B)], 2.0 * _TRI_A
    else:
        pts, qw = [(_TRI_A, _TRI_A, _TRI_A)], 2.0
    sel = ((iBCB[:, 0] & 2) != 0) | np.isin(iBCB[:, 1], list(dirichlet_srfs))
    for e in np.nonzero(sel)[0]:
        g = ienb[e, :3] - 1
        v1, v2 = x[g[1]] - x[g[0]], x[g[2]] - x[g[0]]
        t = np.array([v1[1] * v2[2] - v2[1] * v1[2], v2[0] * v1[2] - v1[0] * v2[2], v1[0] * v2[1] - v2[0] * v1[1]])
        temp = 1.0 / np.sqrt(t[0] ** 2 + t[1] ** 2 + t[2] ** 2)
        bn = t * temp
        W = qw / (4.0 * temp)
        for q in pts:
            for a in range(3):
                nabi[g[a]] += q[a] * (bn * W)
                nasc[g[a]] += q[a] * W
    return nabi, nasc


class RCR:
    """The numerical RCR of a coupled surface (RCR.cxx): Rp, C, Rd and the
    distal pressure table pdist (time, value), linear in time and constant
    beyond its ends; P, Q the surface's pressure and flow at t_n."""

    def __init__(self, Rp, C, Rd, pdist, P0, Q0):
        self.Rp, self.C, self.Rd = float(Rp), float(C), float(Rd)
        pdist = np.asarray(pdist, dtype=np.float64)
        self.pt, self.pv = pdist[:, 0], pdist[:, 1]
        self.P, self.Q = float(P0), float(Q0)

    def pd(self, t):
        return float(np.interp(t, self.pt, self.pv))

    def coefficients(self, t0, t1, D):
        k = self.C * self.Rd / D
        den = 1.0 + k
        c1 = (self.Rd + self.Rp * (1.0 + k)) / den
        c2 = (k * (self.P + self.pd(t1) - self.pd(t0) - self.Rp * self.Q) + self.pd(t1)) / den
        return c1, c2


class PhastaCase:
    """A case's arrays (as read by phasta.read_geombc / read_restart, in the
    solver's column order) stepped on the GPU.

    x(numnp,3), ien(nel,4), ienb(nelb,4), iBCB(nelb,2), BCB(nelb,3,ndBCB),
    iBC(nshg), y, ac (nshg,4), u, uref (nshg,3), nodetag(numnp), ndsurf
    (nshg, the "global node surface number"); bct_records from
    bct.read_bct; rcr = {srfID: (Rp, C, Rd, pdist)}; rho, mu, Delt,
    rhoinf; wall a WallParams (None: rigid walls); ls the memLS NS
    settings (Kspace, tolerances, limits); dirichlet_srfs the surfaces
    NABI also covers."""

    def __init__(self, x, ien, ienb, iBCB, BCB, iBC, y, ac, u, uref, nodetag, ndsurf, bct_records, rcr,
                 rho, mu, Delt, rhoinf, wall=None, ls=None, nonlinear=6, res_criteria=0.005, min_iters=2,
                 rescontrol=True, influx=0.5, nqb=3, dirichlet_srfs=(), bct_time_scale=1.0):
        import torch
        self.torch = torch
        dev = self.dev = torch.device("cuda:0")
        self.x = np.asarray(x, dtype=np.float64)
        N = self.N = self.x.shape[0]
        self.iBC = np.asarray(iBC, dtype=np.int32)
        self.ndsurf = np.asarray(ndsurf)
        self.params = p = form_params(rho=rho, mu=mu, delt=Delt, rhoinf=rhoinf, ideformwall=1 if wall else 0)
        self.Delt = float(Delt)
        self.nonlinear, self.res_criteria, self.min_iters = int(nonlinear), float(res_criteria), int(min_iters)
        self.rescontrol = bool(rescontrol)
        self.colm, self.rowp = genadj(ien, N)
        self.form = f = ElmGMRInterior(ien, self.colm, self.rowp, N, iBC=self.iBC, params=p)
        f.set_boundary(ienb, iBCB, BCB, nqb=nqb, influx=influx)
        self.wall = wall is not None
        T = self.T
        self.dx = T(self.x)
        self.y, self.ac = T(y), T(ac)
        self.u = T(u)
        if self.wall:
            f.set_wall(wall, nodetag=nodetag, uref=uref)
            self.ua = torch.zeros((3, N), dtype=torch.float64, device=dev)
            f.set_displacement(self.ua)
        self.nabi, self.nasc = surface_integrals(self.x, ienb, iBCB, nqb=nqb, dirichlet_srfs=dirichlet_srfs)
        # memLS: the Dirichlet nodes (sV = 1 for a free component of a
        # partly coded node, itrdrv.f90:934-992), then one Neumann face per
        # coupled surface with its nodes' NABI (AddNeumannBCTomemLS)
        self.srfs = sorted(rcr)
        fixed = np.nonzero((self.iBC >> 3) & 7)[0]
        sV = np.zeros((3, fixed.size), order="F")
        for c in range(3):
            sV[c] = ((self.iBC[fixed] >> (3 + c)) & 1) == 0
        self.lhs = memLS_LHS_CREATE(N, N, self.rowp.size, None, self.colm, self.rowp, 1 + len(self.srfs))
        memLS_BC_CREATE(self.lhs, 1, fixed.size, 3, BC_TYPE_Dir, fixed + 1, sV)
        self.faces = []
        for k, s in enumerate(self.srfs):
            nodes = np.nonzero(self.ndsurf == s)[0]
            memLS_BC_CREATE(self.lhs, 2 + k, nodes.size, 3, BC_TYPE_Neu, nodes + 1,
                            np.asfortranarray(self.nabi[nodes].T))
            self.faces.append((torch.from_numpy(nodes.astype(np.int32)).to(dev),
                               torch.from_numpy(np.ascontiguousarray(self.nabi[nodes])).to(dev)))
        if ls is None:
            ls = dict(relTol=0.4, dimKry=45, maxItr=10, relTolIn=(0.01, 0.2), maxItrIn=(1, 200))
        self.ls = memLS_LS_CREATE(LS_TYPE_NS, relTol=ls["relTol"], dimKry=ls["dimKry"], maxItr=ls["maxItr"],
                                  relTolIn=list(ls["relTolIn"]), maxItrIn=list(ls["maxItrIn"]))
        self.bct = _bct.BCt(bct_records, self.x, self.iBC, time_scale=bct_time_scale)
        self.dibc = torch.from_numpy(self.iBC).to(dev)
        self.res4 = torch.zeros((N, 4), dtype=torch.float64, device=dev)
        self.dq = torch.zeros(1, dtype=torch.float64, device=dev)
        y0 = np.asarray(y, dtype=np.float64)
        self.rcr = {s: RCR(*rcr[s], P0=self.mean_pressure_host(y0, s), Q0=self.flow_host(y0, s)) for s in self.srfs}
        self.n = 0

    def T(self, a):
        """Host (n, k) array -> device (k, n) C tensor (its Fortran layout)."""
        return self.torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64, order="F").T)).to(self.dev)

    def flow_host(self, y, srf):
        nodes = np.nonzero(self.ndsurf == srf)[0]
        return float(np.sum(self.nabi[nodes] * np.asarray(y)[nodes, :3]))

    def mean_pressure_host(self, y, srf):
        nodes = np.nonzero(self.ndsurf == srf)[0]
        return float(np.sum(self.nasc[nodes] * np.asarray(y)[nodes, 3]) / np.sum(self.nasc[nodes]))

    def flow(self, srf):
        """GetFlowQ of the current state (sum NABI . u over the surface)."""
        nodes = np.nonzero(self.ndsurf == srf)[0]
        yn = self.y[:3, nodes].cpu().numpy().T
        return float(np.sum(self.nabi[nodes] * yn))

    def mean_pressure(self, srf):
        nodes = np.nonzero(self.ndsurf == srf)[0]
        pn = self.y[3, nodes].cpu().numpy()
        return float(np.sum(self.nasc[nodes] * pn) / np.sum(self.nasc[nodes]))

    def step(self):
        """One timestep; returns its record (flows, pressures, iterations)."""
        torch, f, p = self.torch, self.form, self.params
        n, dt = self.n, self.Delt
        BC = np.zeros((self.N, 7), order="F")
        self.bct.apply((n + 1) * dt, BC)
        dBC = self.T(BC)
        cs, cu = {}, {}
        for s in self.srfs:
            m = self.rcr[s]
            cs[s] = m.coefficients(dt * n, dt * (n + p.alfi), p.alfi * dt)
            cu[s] = m.coefficients(dt * n, dt * (n + 1.0), dt)
        faceres = np.array([0.0] + [cs[s][0] * p.alfi * p.gami * dt for s in self.srfs])
        yold, acold, uold = self.y, self.ac, self.u
        y, ac = torch.empty_like(yold), torch.empty_like(acold)
        ya, aca = torch.empty_like(yold), torch.empty_like(acold)
        f.predict(yold, acold, y, ac)
        u = uold.clone()
        if self.wall:
            f.wall_predict(uold, yold, acold, u)
        f.itr_bc(self.dibc, dBC, y)
        its = []
        for it in range(1, self.nonlinear + 1):
            if self.wall:
                f.wall_alpha(uold, u, self.ua)
            f.yalpha(yold, acold, y, ac, ya, aca)
            f.to_lhs(ya, aca, self.dx, self.res4, self.lhs)
            for s, (glob, nabi) in zip(self.srfs, self.faces):
                f.elmpvsq_coef(glob, nabi, cs[s][0], cs[s][1], ya, self.res4, self.dq)
            f.bc3res(self.res4, res4_layout=True)
            f.sync()
            totres = float(torch.sqrt(torch.sum(self.res4 * self.res4) / self.N))
            memLS_SOLVE_dev(self.lhs, self.ls, 4, self.res4, res=faceres)
            its.append((totres, self.ls.RI["itr"], self.ls.GM["itr"], self.ls.CG["itr"]))
            f.correct(y, ac, self.res4)
            if self.wall:
                f.wall_correct(u, self.res4)
            f.itr_bc(self.dibc, dBC, y)
            f.sync()
            if self.rescontrol and totres < self.res_criteria and it >= self.min_iters:
                break
        self.y, self.ac, self.u = y, ac, u
        self.n = n + 1
        rec = {"step": self.n, "iterations": its}
        for s in self.srfs:
            Q1 = self.flow(s)
            m = self.rcr[s]
            m.P = cu[s][0] * Q1 + cu[s][1]
            m.Q = Q1
            rec[s] = {"Q": Q1, "P": m.P, "p_mean": self.mean_pressure(s)}
        return rec

    def restart(self, y, ac, u, step, pressures):
        """Continue from a restart at `step` (currentTimestepIndex): the
        state (nshg,k) arrays, each RCR surface's pressure {srfID: P} (its
        PHistRCR row, RCR.cxx:109-128) and its flow from the state."""
        self.y, self.ac, self.u = self.T(y), self.T(ac), self.T(u)
        self.n = int(step)
        for s in self.srfs:
            self.rcr[s].P = float(pressures[s])
            self.rcr[s].Q = self.flow_host(np.asarray(y), s)

    def state(self):
        """(y, ac, u) on the host, (nshg, k) arrays."""
        return tuple(t.cpu().numpy().T.copy(order="F") for t in (self.y, self.ac, self.u))


def wall_from_solver_inp(**kw):
    """wall_params with the reference basic case's solver.inp values."""
    d = dict(rhovw=0.00106, thicknessvw=0.3, evw=7.0e5, rnuvw=0.5, rshearconstantvw=0.833333,
             tissSuppStiffCoeff=0.10, tissSuppDampCoeff=0.01, tissSuppRingStiffCoeff=5e5,
             tissSuppRingDampCoeff=0.01, iwallmassfactor=1, iwallstiffactor=1, iwallsupp=1, iwalldamp=1,
             iringsupp=1, iringdamp=1)
    d.update(kw)
    return wall_params(**d)
