"""Ensembles of instances (parameter samples on meshes of uneven size)
formed and solved together on one GPU.

The instances' meshes are joined into one union mesh with disjoint node
ranges (a block-diagonal CSR graph), formed by one formation handle whose
elements carry their instance's viscosity (crimson_form_set_instances), so a
single launch sequence forms every instance; that sequence is captured once
in a HIP graph and replayed for each nonlinear iteration.  The memLS matrix
of the union has one Dirichlet face (every instance's fixed nodes) and one
resistance face per instance, so one NS solve advances all instances; its
Krylov space spans all instances, so the instances converge together (the
stopping test is on the union's residual).  Each instance's blocks and
residual rows are exactly those of the instance formed alone.

Not in the reference, which runs one case per process.
"""
from __future__ import annotations

import numpy as np

from . import ElmGMRInterior, form_params, genadj, memLS_BC_CREATE, memLS_LHS_CREATE, BC_TYPE_Dir, BC_TYPE_Neu
from . import mesh as M


def uneven_cylinders(K=48, seed=2024):
    """The batched-ensemble workload: 8 cylinders (radius 1) of seeded
    lengths uniform in 4-20 cm and element counts uniform in 40k-600k tets
    (the hex size h follows from the count: 6 tets per hex, pi/4 of the box
    kept by the clip), and K instances dealt to them in turn with viscosity
    lognormal(ln 0.04, 0.2) and outlet resistance uniform 800-2000.
    Returns (meshes of the instances, the 8 distinct meshes, mu(K), R(K))."""
    rng = np.random.default_rng(seed)
    shapes = []
    for _ in range(8):
        Lc, Tc = rng.uniform(4.0, 20.0), rng.uniform(4e4, 6e5)
        h = (18.85 * Lc / Tc) ** (1.0 / 3.0)
        n = max(4, round(2.0 / h))
        shapes.append(M.cylinder_mesh(n, n, max(4, round(Lc / h)), length=Lc))
    meshes = [shapes[k % 8] for k in range(K)]
    mu = np.exp(rng.normal(np.log(0.04), 0.2, K))
    R = rng.uniform(800.0, 2000.0, K)
    return meshes, shapes, mu, R


class Ensemble:
    def __init__(self, meshes, mu, R, precision=64, params=None, outlet_natural=True, boundary=False):
        self.meshes = list(meshes)
        K = len(self.meshes)
        self.K = K
        self.mu = np.asarray(mu, dtype=np.float64)
        self.R = np.asarray(R, dtype=np.float64)
        if self.mu.shape != (K,) or self.R.shape != (K,):
            raise ValueError("one viscosity and one resistance per instance")
        self.node_off = np.concatenate([[0], np.cumsum([m.numnp for m in self.meshes])]).astype(np.int64)
        self.elem_off = np.concatenate([[0], np.cumsum([m.nel for m in self.meshes])]).astype(np.int64)
        N = int(self.node_off[-1])
        self.N = N
        self.ien = np.asfortranarray(np.concatenate([m.ien + int(o) for m, o in zip(self.meshes, self.node_off)])
                                     .astype(np.int32))
        self.x = np.asfortranarray(np.concatenate([m.x for m in self.meshes]))
        inst = np.concatenate([np.full(m.nel, k, np.uint16) for k, m in enumerate(self.meshes)])
        ibc, fixed, outs, nabis = [], [], [], []
        for k, m in enumerate(self.meshes):
            b, _ = M.dirichlet_codes(m, outlet_natural=outlet_natural)
            ibc.append(b)
            o = int(self.node_off[k])
            fixed.append(np.nonzero(b)[0] + o)
            out = np.nonzero(m.outlet & (b == 0))[0]
            outs.append(out + o)
            nabis.append(m.nabi()[out])
        self.iBC = np.concatenate(ibc).astype(np.int32)
        self.fixed = np.concatenate(fixed)
        self.outlets = outs
        self.nabi = nabis
        self.colm, self.rowp = genadj(self.ien, N)
        p = params if params is not None else form_params(precision=precision)
        p.precision = precision
        self.params = p
        self.form = ElmGMRInterior(self.ien, self.colm, self.rowp, N, iBC=self.iBC, params=p)
        self.inst = inst
        self.form.set_instances(inst, self.mu)
        self.nelb = 0
        if boundary:
            # every instance's boundary elements (inlet, wall, outlet), in
            # instance order, on the union's node numbers
            cache = {}
            for m in self.meshes:  # instances may share a mesh object
                if id(m) not in cache:
                    cache[id(m)] = M.cylinder_boundary(m)
            parts = [cache[id(m)] for m in self.meshes]
            ienb = np.concatenate([b[0] + int(o) for b, o in zip(parts, self.node_off)]).astype(np.int32)
            iBCB = np.concatenate([b[1] for b in parts]).astype(np.int32)
            BCB = np.concatenate([b[2] for b in parts])
            self.form.set_boundary(ienb, iBCB, BCB)
            self.nelb = ienb.shape[0]
        self.lhs = memLS_LHS_CREATE(N, N, self.rowp.size, None, self.colm, self.rowp, 1 + K)
        memLS_BC_CREATE(self.lhs, 1, self.fixed.size, 3, BC_TYPE_Dir, self.fixed + 1,
                        np.zeros((3, self.fixed.size), order="F"))
        for k in range(K):
            memLS_BC_CREATE(self.lhs, 2 + k, self.outlets[k].size, 3, BC_TYPE_Neu, self.outlets[k] + 1,
                            np.asfortranarray(self.nabi[k].T))
        self.graph = None

    def face_res(self):
        """memLS face coupling coefficients: 0 on the Dirichlet face, R_k
        alpha_f gamma Delt on instance k's outlet (elmgmr.f90:1806)."""
        p = self.params
        return np.concatenate([[0.0], self.R * p.alfi * p.gami * p.Delt])

    def capture(self, y, ac, x, res4, autotune=True):
        """Capture the union formation on these device arrays (torch
        tensors in the C ABI layouts) into a HIP graph, after timing its
        launch configurations on them and keeping the fastest
        (ElmGMRInterior.autotune; self.tuned holds the choice and times)."""
        self.tuned = self.form.autotune(y, ac, x, res4, self.lhs) if autotune else None
        self.graph = self.form.graph_to_lhs(y, ac, x, res4, self.lhs)
        return self.graph

    def set_parameters(self, mu=None, R=None):
        """New instance viscosities and/or outlet resistances (the union's
        graph and matrix are unchanged)."""
        if mu is not None:
            self.mu = np.asarray(mu, dtype=np.float64)
            self.form.set_instances(self.inst, self.mu)
        if R is not None:
            self.R = np.asarray(R, dtype=np.float64)

    def rows(self, k):
        return slice(int(self.node_off[k]), int(self.node_off[k + 1]))

    def nnz_range(self, k):
        return slice(int(self.colm[self.node_off[k]] - 1), int(self.colm[self.node_off[k + 1]] - 1))
