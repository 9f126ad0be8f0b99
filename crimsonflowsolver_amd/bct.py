"""Time-varying Dirichlet velocities from ``bct.dat`` (host side, numpy).

The reference reads ``bct.dat`` once (``initBCt``, bctint.f90:240-333) and,
at every timestep, interpolates it into the BC array (``BCint``,
bctint.f90:335-389):

* file: a first line ``ntv nptsmax``, then for each of the ntv nodes a line
  ``x1 x2 x3 ntpts`` followed by ntpts lines ``u1 u2 u3 t``;
* each record is attached to the first node in numbering order whose iBC
  has all three velocity bits (``ibits(iBC,3,3) == 7``) and whose squared
  distance to (x1, x2, x3) is below ``epsd`` (1e-6); a record that matches
  no node is skipped; times are multiplied by "BCT Time Scale Factor";
* at time ``timel`` every node's BC(:,3:5) is first set to the mesh velocity
  (zero without ALE), then each attached node's to the linear interpolation
  of its series at ``t = timel - nper T`` (T its last time, nper the whole
  periods completed), between the first sample with time above t and the
  one before it; when no later sample exists the node keeps the mesh
  velocity.
"""
from __future__ import annotations

import numpy as np

EPSD = 1.0e-6  # squared distance (bctint.f90: "new distance to avoid problems when using mm")


def read_bct(path):
    """[(x(3), table(ntpts, 4) with columns u1 u2 u3 t)] in file order."""
    with open(path) as f:
        first = f.readline().split()
        ntv = int(first[0])
        out = []
        for _ in range(ntv):
            head = f.readline().split()
            if len(head) < 4:
                raise ValueError(f"{path}: truncated record header")
            x = np.array([float(v) for v in head[:3]])
            n = int(head[3])
            rows = [f.readline().split()[:4] for _ in range(n)]
            if any(len(r) < 4 for r in rows):
                raise ValueError(f"{path}: record of {n} samples is short")
            tab = np.array(rows, dtype=np.float64).reshape(n, 4)
            out.append((x, tab))
    return out


class BCt:
    """The attached series of a bct.dat on a mesh (initBCt)."""

    def __init__(self, records, x, iBC, time_scale=1.0, epsd=EPSD):
        x = np.asarray(x, dtype=np.float64)
        iBC = np.asarray(iBC)
        cand = np.nonzero(((iBC >> 3) & 7) == 7)[0]
        nodes, tabs = [], []
        for xr, tab in records:
            d = x[cand] - xr
            dd = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
            hit = np.nonzero(dd < epsd)[0]
            if hit.size == 0:
                continue  # on another part (the reference reads past it)
            t = tab.copy()
            t[:, 3] = t[:, 3] * time_scale
            nodes.append(int(cand[hit[0]]))
            tabs.append(t)
        self.nodes = np.array(nodes, dtype=np.int64)
        self.tables = tabs

    def apply(self, timel, BC, mesh_velocity=(0.0, 0.0, 0.0)):
        """BCint: BC(nshg, ndofBC) (Fortran columns 3:5 = 0-based 2:5) at
        time timel, in place; returns BC."""
        BC[:, 2:5] = np.asarray(mesh_velocity, dtype=np.float64)
        for node, tab in zip(self.nodes, self.tables):
            T = tab[-1, 3]
            nper = int(timel / T)  # Fortran's conversion to integer truncates
            t = timel - nper * T
            later = np.nonzero(tab[1:, 3] > t)[0]
            if later.size == 0:
                continue
            j = later[0] + 1
            wr = (t - tab[j - 1, 3]) / (tab[j, 3] - tab[j - 1, 3])
            BC[node, 2:5] = tab[j - 1, :3] * (1.0 - wr) + tab[j, :3] * wr + np.asarray(mesh_velocity)
        return BC
