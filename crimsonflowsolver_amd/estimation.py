"""Parameter estimation client: the reduced-order UKF (roukf.ROUKF) over a
flow problem whose sigma points run as one ensemble (ensemble.Ensemble) on
the GPU, the role the reference's CrimsonVerdandiModel plays for its
ROUKFModified (estimation/src/CrimsonVerdandiModel.cxx:711-828: Forward
advances every sigma point's flow state one timestep through itrdrv;
ApplyOperator; the observation manager compares nodal values).

The estimated parameters are the viscosity and/or the outlet resistance of
a cylinder with a prescribed inflow; the state of each sigma point is its
(y, ac) fields plus the parameters.  Each step: p + 1 sigma states are
loaded into the union's node ranges, one generalized-alpha timestep
(itrPredict, itrBC, then per nonlinear iteration itrYAlpha, the formation
into the union's memLS matrix, bc3Res, ElmpvsQ's resistance term of each
instance, one memLS NS solve, itrCorrect, itrBC) advances all of them
together, and the ROUKF analysis uses the pressures at the observed nodes.
The sigma states pass through the host between steps (the p x p algebra
is the host's, as in the reference).
"""
from __future__ import annotations

import numpy as np

from . import form_params, memLS_LS_CREATE, memLS_SOLVE_dev, LS_TYPE_NS
from . import mesh as M
from .ensemble import Ensemble
from .roukf import ROUKF

PARAMS = ("mu", "R")


class CylinderFlow:
    """p + 1 copies of `mesh` as one ensemble; `advance(Y, AC, mu, R)` runs
    one timestep of every copy (Y, AC: (K, N, 4) numpy fields, mu, R: (K,))
    and returns the new fields."""

    def __init__(self, mesh, K, inflow, precision=64, niter=3, boundary=True, solver=None):
        import torch
        self.torch = torch
        self.m, self.K, self.niter = mesh, K, niter
        self.ens = Ensemble([mesh] * K, np.full(K, 0.04), np.full(K, 1000.0), precision=precision,
                            boundary=boundary)
        N1 = mesh.numnp
        self.N1 = N1
        dev = torch.device("cuda:0")
        self.dev = dev
        p = self.ens.params
        self.p = p
        ibc1, _ = M.dirichlet_codes(mesh)
        rr = mesh.x[:, 0] ** 2 + mesh.x[:, 1] ** 2
        inl = mesh.inlet & ~mesh.wall
        bc1 = np.zeros((N1, 7), order="F")
        bc1[inl, 4] = 2.0 * inflow * np.clip(1.0 - rr[inl], 0.0, None)
        bc = np.asfortranarray(np.concatenate([bc1] * K))
        T = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).to(dev)
        self.T = T
        self.dx = T(self.ens.x)
        self.dBC = T(bc)
        self.dibc = torch.from_numpy(self.ens.iBC.astype(np.int32)).to(dev)
        self.outs = [(torch.from_numpy(o.astype(np.int32)).to(dev),
                      torch.from_numpy(np.ascontiguousarray(nb)).to(dev)) for o, nb in
                     zip(self.ens.outlets, self.ens.nabi)]
        self.dq = torch.zeros(1, dtype=torch.float64, device=dev)
        self.r4 = torch.zeros((self.ens.N, 4), dtype=torch.float64, device=dev)
        self.kw = solver or dict(relTol=1e-6, dimKry=60, maxItr=10, relTolIn=[1e-4, 1e-4], maxItrIn=[10, 500])

    def advance(self, Y, AC, mu, R):
        torch = self.torch
        ens, f = self.ens, self.ens.form
        ens.set_parameters(mu=mu, R=R)
        yold = self.T(np.concatenate(list(Y)))
        acold = self.T(np.concatenate(list(AC)))
        yc, acc = torch.empty_like(yold), torch.empty_like(acold)
        ya, aca = torch.empty_like(yold), torch.empty_like(acold)
        f.predict(yold, acold, yc, acc)
        f.itr_bc(self.dibc, self.dBC, yc)
        ls = memLS_LS_CREATE(LS_TYPE_NS, **self.kw)
        for _ in range(self.niter):
            f.yalpha(yold, acold, yc, acc, ya, aca)
            f.to_lhs(ya, aca, self.dx, self.r4, ens.lhs)
            f.bc3res(self.r4, res4_layout=True)
            for k, (g, nb) in enumerate(self.outs):
                f.elmpvsq(g, nb, float(ens.R[k]), ya, self.r4, self.dq)
            f.sync()
            memLS_SOLVE_dev(ens.lhs, ls, 4, self.r4, res=ens.face_res())
            f.correct(yc, acc, self.r4)
            f.itr_bc(self.dibc, self.dBC, yc)
        f.sync()
        yn = yc.cpu().numpy().T.reshape(self.K, self.N1, 4)
        an = acc.cpu().numpy().T.reshape(self.K, self.N1, 4)
        return yn, an


class Estimator:
    """ROUKF over `names` (a subset of ("mu", "R")) of a CylinderFlow.
    theta0, P0: prior mean and covariance; fixed: values of the parameters
    not estimated; obs_nodes: nodes whose pressure is observed, with error
    variance sigma2 each."""

    def __init__(self, mesh, names, theta0, P0, fixed, y0, ac0, obs_nodes, sigma2, inflow, **kw):
        self.names = tuple(names)
        if not set(self.names) <= set(PARAMS) or not self.names:
            raise ValueError(f"estimated parameters must be among {PARAMS}")
        p = len(self.names)
        self.flow = CylinderFlow(mesh, p + 1, inflow, **kw)
        self.fixed = dict(fixed)
        self.obs = np.asarray(obs_nodes)
        self.nf = 8 * mesh.numnp
        x0 = np.concatenate([np.asarray(y0).ravel(), np.asarray(ac0).ravel(), np.asarray(theta0, dtype=float)])
        L0 = np.zeros((x0.size, p))
        L0[self.nf:, :] = np.eye(p)
        self.kf = ROUKF(x0, L0, np.atleast_2d(P0), np.eye(self.obs.size) * sigma2)

    def _unpack(self, X):
        N = self.flow.N1
        Y = [X[:4 * N, i].reshape(N, 4) for i in range(X.shape[1])]
        AC = [X[4 * N:8 * N, i].reshape(N, 4) for i in range(X.shape[1])]
        th = X[self.nf:, :]
        par = {n: th[j] for j, n in enumerate(self.names)}
        mu = par.get("mu", np.full(X.shape[1], self.fixed.get("mu", 0.04)))
        R = par.get("R", np.full(X.shape[1], self.fixed.get("R", 1000.0)))
        return Y, AC, mu, R

    def step(self, y_obs):
        """One forward step of every sigma point, then the analysis with the
        observed pressures y_obs; returns the parameter estimate."""
        def advance(X):
            Y, AC, mu, R = self._unpack(X)
            yn, an = self.flow.advance(Y, AC, mu, R)
            return np.vstack([yn.reshape(X.shape[1], -1).T, an.reshape(X.shape[1], -1).T, X[self.nf:, :]])

        self.kf.forward(advance)
        N = self.flow.N1
        self.kf.analyze(lambda X: np.asarray(y_obs)[:, None] - X[:4 * N].reshape(N, 4, -1)[self.obs, 3, :])
        return self.theta

    @property
    def theta(self):
        return self.kf.x[self.nf:].copy()
