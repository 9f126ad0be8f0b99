"""Seeded synthetic cylinder meshes and flow states (host side, numpy).

The meshes are the structured-hex cylinders of the benchmark workloads: an
nx x ny x nz hex grid over [-R, R]^2 x [0, L], hexes kept when their centroid
radius is <= R, each kept hex split into 6 P1 tets by the Kuhn (Freudenthal)
split along its (0,0,0)-(1,1,1) diagonal, which is conforming across hexes.

Arrays follow the reference's conventions (flowsolver/src/common.f90 and the
PHASTA geombc layout read by readnblk.f90): coordinates x(numnp,3) and nodal
fields y(nshg,4) with columns (u1,u2,u3,p) are column-major; the element
connectivity ien(nel,4) holds 1-based node numbers.  Elements are oriented so
that det[x1-x4, x2-x4, x3-x4] > 0, the orientation for which the reference's
e3metric (e3metric.f90:42-67) yields a positive WdetJ.

Boundary-condition codes use the reference's iBC bits (itrbc.f90, bc3res.f90):
bits 3,4,5 constrain velocity components 1,2,3.
"""
from __future__ import annotations

import dataclasses
import itertools

import numpy as np

IBC_VEL = (1 << 3) | (1 << 4) | (1 << 5)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n uint64 values of the splitmix64 sequence started at `seed`."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform_pm1(seed: int, n: int) -> np.ndarray:
    """n doubles uniform in [-1, 1) from splitmix64(seed)."""
    z = splitmix64(seed, n)
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


@dataclasses.dataclass
class CylinderMesh:
    x: np.ndarray          # (numnp, 3) float64, Fortran order
    ien: np.ndarray        # (nel, 4) int32, 1-based, Fortran order
    inlet: np.ndarray      # (numnp,) bool
    outlet: np.ndarray     # (numnp,) bool
    wall: np.ndarray       # (numnp,) bool
    outlet_tris: np.ndarray  # (ntri, 3) int32 0-based, outward normal +z
    grid: tuple
    radius: float
    length: float

    @property
    def numnp(self) -> int:
        return self.x.shape[0]

    @property
    def nel(self) -> int:
        return self.ien.shape[0]

    def nabi(self) -> np.ndarray:
        """NABI(i,1:3) = integral of N_i n over the outlet face (the quantity
        pvsqbi.f90 forms for a resistance surface); exact for P1: each
        triangle adds area/3 * n to each of its nodes."""
        out = np.zeros((self.numnp, 3))
        t = self.outlet_tris
        if t.size:
            p0, p1, p2 = self.x[t[:, 0]], self.x[t[:, 1]], self.x[t[:, 2]]
            an = 0.5 * np.cross(p1 - p0, p2 - p0)  # area * normal
            for k in range(3):
                np.add.at(out, t[:, k], an / 3.0)
        return np.asfortranarray(out)


def cylinder_mesh(nx: int, ny: int, nz: int, radius: float = 1.0,
                  length: float = 10.0, clip: bool = True) -> CylinderMesh:
    """Structured-hex cylinder split into Kuhn tets (see module docstring)."""
    hx, hy, hz = 2.0 * radius / nx, 2.0 * radius / ny, length / nz
    ci = (np.arange(nx) + 0.5) * hx - radius
    cj = (np.arange(ny) + 0.5) * hy - radius
    cell2d = np.ones((nx, ny), bool)
    if clip:
        cell2d = (ci[:, None] ** 2 + cj[None, :] ** 2) <= radius * radius
    I, J = np.nonzero(cell2d)                      # kept columns
    ncol = I.size
    K = np.arange(nz)
    # hexes ordered k-slowest, then by column order (j, i lexicographic)
    order = np.lexsort((I, J))
    I, J = I[order], J[order]
    hi = np.tile(I, nz)
    hj = np.tile(J, nz)
    hk = np.repeat(K, ncol)
    NX, NY = nx + 1, ny + 1

    def lex(i, j, k):
        return i + NX * (j + NY * k)

    tets = []
    for perm in itertools.permutations(range(3)):
        v = [np.zeros(3, int)]
        cur = np.zeros(3, int)
        for d in perm:
            cur = cur.copy()
            cur[d] += 1
            v.append(cur)
        tets.append(np.stack([lex(hi + o[0], hj + o[1], hk + o[2]) for o in v], axis=1))
    # element order: hex, then the 6 tets of the hex
    T = np.stack(tets, axis=1).reshape(-1, 4)
    used = np.unique(T)
    remap = np.full(NX * NY * (nz + 1), -1, np.int64)
    remap[used] = np.arange(used.size)
    T = remap[T]
    li = used % NX
    lj = (used // NX) % NY
    lk = used // (NX * NY)
    x = np.stack([li * hx - radius, lj * hy - radius, lk * hz], axis=1)
    # orientation: det[x1-x4, x2-x4, x3-x4] > 0
    d1 = x[T[:, 0]] - x[T[:, 3]]
    d2 = x[T[:, 1]] - x[T[:, 3]]
    d3 = x[T[:, 2]] - x[T[:, 3]]
    det = np.einsum("ij,ij->i", d1, np.cross(d2, d3))
    neg = det < 0
    T[neg, 0], T[neg, 1] = T[neg, 1].copy(), T[neg, 0].copy()
    # boundary tagging
    inlet = lk == 0
    outlet = lk == nz
    pad = np.zeros((nx + 2, ny + 2), bool)
    pad[1:-1, 1:-1] = cell2d
    # node (i,j) touches cells (i-1..i, j-1..j) -> pad indices (i..i+1, j..j+1)
    allkept = pad[li, lj] & pad[li + 1, lj] & pad[li, lj + 1] & pad[li + 1, lj + 1]
    wall = ~allkept
    # outlet triangles: tet faces with all three nodes on the outlet plane
    faces = np.concatenate([T[:, [0, 1, 2]], T[:, [0, 1, 3]], T[:, [0, 2, 3]], T[:, [1, 2, 3]]])
    on = outlet[faces].all(axis=1)
    tri = faces[on]
    if tri.size:
        p0, p1, p2 = x[tri[:, 0]], x[tri[:, 1]], x[tri[:, 2]]
        nzc = np.cross(p1 - p0, p2 - p0)[:, 2]
        flip = nzc < 0
        tri[flip, 1], tri[flip, 2] = tri[flip, 2].copy(), tri[flip, 1].copy()
    return CylinderMesh(x=np.asfortranarray(x), ien=np.asfortranarray((T + 1).astype(np.int32)),
                        inlet=inlet, outlet=outlet, wall=wall,
                        outlet_tris=tri.astype(np.int32), grid=(nx, ny, nz),
                        radius=radius, length=length)


def poiseuille_state(mesh: CylinderMesh, peak: float = 30.0, dp: float = 1330.0,
                     noise: float = 0.05, seed: int = 20240611):
    """Seeded flow state: u_z = peak (1 - r^2/R^2) (clamped at 0) plus
    +/- noise*peak uniform additive perturbation on every velocity component,
    p = dp (1 - z/L), ac = 0.  Returns y(nshg,4) (u1,u2,u3,p), ac(nshg,4)."""
    n = mesh.numnp
    x = mesh.x
    r2 = (x[:, 0] ** 2 + x[:, 1] ** 2) / mesh.radius ** 2
    xi = uniform_pm1(seed, 3 * n).reshape(3, n)
    y = np.zeros((n, 4), order="F")
    y[:, 0] = noise * peak * xi[0]
    y[:, 1] = noise * peak * xi[1]
    y[:, 2] = peak * np.clip(1.0 - r2, 0.0, None) + noise * peak * xi[2]
    y[:, 3] = dp * (1.0 - x[:, 2] / mesh.length)
    ac = np.zeros((n, 4), order="F")
    return y, ac


def dirichlet_codes(mesh: CylinderMesh, outlet_natural: bool = True):
    """iBC with all three velocity bits on inlet and wall nodes; BC(nshg,7)
    holding the Dirichlet values in columns 3..5 (1-based, itrbc.f90:83-86)
    set to zero (callers fill the inlet profile)."""
    n = mesh.numnp
    ibc = np.zeros(n, np.int32)
    fixed = mesh.inlet | mesh.wall
    if not outlet_natural:
        fixed |= mesh.outlet
    ibc[fixed] = IBC_VEL
    bc = np.zeros((n, 7), order="F")
    return ibc, bc


def csr_graph(ien0: np.ndarray, nshg: int):
    """Node-adjacency CSR including self (what genadj.f90 builds), 0-based:
    returns (rowptr[nshg+1], cols[nnz]) with ascending columns per row."""
    e = np.asarray(ien0, dtype=np.int64)
    a = np.repeat(e, 4, axis=1).ravel()
    b = np.tile(e, (1, 4)).ravel()
    key = np.unique(a * nshg + b)
    rows = key // nshg
    cols = (key % nshg).astype(np.int32)
    rowptr = np.zeros(nshg + 1, np.int64)
    np.add.at(rowptr, rows + 1, 1)
    return np.cumsum(rowptr).astype(np.int32), cols
