"""Pins of the oracle's deformable-wall terms (cfo_boundary_wall: e3bvar's
membrane, mass, tissue support, damping and ring tethering, e3b's use of
them, and the wall tangent through bc3lhs and fillsparseI), on the CPU.

* membrane patch test: a flat plate of wall faces (alternating diagonals,
  each face's own first edge as its local x axis) under a uniform in-plane
  strain carries no force at its interior nodes, and at its edge nodes the
  plane-stress traction sigma n of E/(1-nu^2) [1 nu 0; nu 1 0; 0 0 (1-nu)/2]
  integrated against N_a along the edges (h times); a linear transverse
  displacement gives the transverse shear kappa E/(1-nu^2)(1-nu)/2 the same
  way;
* objectivity: the plate and its displacement rotated by a random rotation
  give the rotated forces;
* the reference displacement adds to the displacement in the strain;
* mass, tissue support and damping: rho_w h a, k u, c v integrated against
  N_a (one third of each face's area for uniform fields), the damping
  velocity also giving the continuity term -N_a u.n;
* ring tethering: for two ring nodes, the edge integral of a linear spring
  along the in-face direction normal to the ring edge, k L/6 (2 u1 + u2).d d;
* the wall tangent against minus the derivative of the residual along the
  increment's path (u by alfi betai dt^2, ac by almi, y by alfi gami dt)
  for random increments, by central differences of the residual (exact for
  these linear terms up to rounding).  e3bvar.f90:752-757 writes the
  membrane block of residual component i and displacement component k at
  LHSwall(3(i-1)+k), where e3lhs's blocks hold [equation i][variable j] at
  lhsK(i + 3(j-1)): the wall blocks are each other's transposes, so the
  assembled tangent matches the derivative with its 3x3 blocks read
  transposed (and misses it read as the interior's blocks: the membrane's
  off-diagonal node blocks are not symmetric).  With ring damping the
  residual takes the first ring node's velocity at both ends while the
  tangent is symmetric in them: that case differs either way.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _plate(nx=6, ny=4, Lx=3.0, Ly=2.0, z0=1.0):
    """Wall faces of a flat plate z = z0 (each split with alternating
    diagonals, nodes ordered so that faces start at different corners) and
    one apex node below each face (the tet's fourth node)."""
    xs = np.linspace(0.0, Lx, nx + 1)
    ys = np.linspace(0.0, Ly, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    pts = [np.stack([X.ravel(), Y.ravel(), np.full(X.size, z0)], axis=1)]
    nid = lambda i, j: i * (ny + 1) + j
    faces = []
    for i in range(nx):
        for j in range(ny):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            if (i + j) % 2:
                tris = [(a, b, c), (a, c, d)]
            else:
                tris = [(b, d, a), (c, d, b)]
            for k, t in enumerate(tris):
                faces.append(t[k:] + t[:k])  # rotate the start corner
    faces = np.array(faces)
    npl = pts[0].shape[0]
    apex = pts[0][faces].mean(axis=1) - np.array([0.0, 0.0, 0.4])
    x = np.concatenate([pts[0], apex])
    ienb = np.concatenate([faces, (npl + np.arange(len(faces)))[:, None]], axis=1) + 1
    return x, np.asfortranarray(ienb, dtype=np.int32), npl, (xs, ys)


def _bcb(nelb, code=16):
    iBCB = np.zeros((nelb, 2), np.int32, order="F")
    iBCB[:, 0] = code
    iBCB[:, 1] = 1
    return iBCB, np.zeros((nelb, 3, 6), order="F")


def _res(x, ienb, y, ac, u, wall, **kw):
    nshg = x.shape[0]
    iBCB, BCB = _bcb(ienb.shape[0], kw.pop("code", 16))
    res = np.zeros((nshg, 4), order="F")
    return O.boundary_wall(ienb, iBCB, BCB, y, ac, u, x, res, wall, **kw)


def _membrane_only(p=None, **kw):
    d = dict(iwallsupp=0, iwalldamp=0, iringsupp=0, iringdamp=0)
    d.update(kw)
    return O.make_wall(p, **d)


def _edge_load(xs, ys, node_of, load):
    """sum over boundary edges of (l/2) load(n) at each end node: the
    integral of N_a times a uniform traction along the plate's edges."""
    f = {}
    edges = []
    for k in range(len(xs) - 1):
        edges.append(((k, 0), (k + 1, 0), (0.0, -1.0), xs[k + 1] - xs[k]))
        edges.append(((k, len(ys) - 1), (k + 1, len(ys) - 1), (0.0, 1.0), xs[k + 1] - xs[k]))
    for k in range(len(ys) - 1):
        edges.append(((0, k), (0, k + 1), (-1.0, 0.0), ys[k + 1] - ys[k]))
        edges.append(((len(xs) - 1, k), (len(xs) - 1, k + 1), (1.0, 0.0), ys[k + 1] - ys[k]))
    for a, b, n, l in edges:
        t = load(np.array(n)) * l / 2
        for e in (a, b):
            f[node_of(*e)] = f.get(node_of(*e), 0.0) + t
    return f


def test_membrane_patch_in_plane_and_transverse():
    x, ienb, npl, (xs, ys) = _plate()
    w = _membrane_only()
    E, nu, h, kap = w.evw, w.rnuvw, w.thicknessvw, w.rshearconstantvw
    G = np.array([[1.3e-3, -0.4e-3], [0.7e-3, -0.9e-3]])
    c = np.array([2.0e-3, -1.1e-3])
    u = np.zeros((x.shape[0], 3))
    u[:, :2] = x[:, :2] @ G.T
    u[:, 2] = x[:, :2] @ c
    z = np.zeros((x.shape[0], 4))
    r = _res(x, ienb, z, z, u, w)
    eps = np.array([G[0, 0], G[1, 1], G[0, 1] + G[1, 0]])
    k = E / (1 - nu * nu)
    sig = k * np.array([[eps[0] + nu * eps[1], 0.5 * (1 - nu) * eps[2]],
                        [0.5 * (1 - nu) * eps[2], eps[1] + nu * eps[0]]])
    q = kap * k * 0.5 * (1 - nu) * c  # transverse shear stress (xz, yz)
    load = lambda n: h * np.array([sig[0] @ n, sig[1] @ n, q @ n])
    node_of = lambda i, j: i * (len(ys)) + j
    want = np.zeros((x.shape[0], 3))
    for nd, f in _edge_load(xs, ys, node_of, load).items():
        want[nd] = -f
    scale = np.abs(want).max()
    assert scale > 0
    err = np.abs(r[:, :3] - want).max() / scale
    print("patch test error", err)
    assert err < 1e-12
    assert np.all(r[:, 3] == 0.0)


def test_objective_under_rotation_and_reference_displacement():
    x, ienb, npl, _ = _plate(4, 3)
    w = _membrane_only()
    rng = np.random.default_rng(3)
    u = rng.normal(size=(x.shape[0], 3)) * 1e-3
    z = np.zeros((x.shape[0], 4))
    r0 = _res(x, ienb, z, z, u, w)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    shift = np.array([0.3, -1.2, 2.0])
    xr = x @ Q.T + shift
    r1 = _res(xr, ienb, z, z, u @ Q.T, w)
    s = np.abs(r0).max()
    assert np.abs(r1[:, :3] - r0[:, :3] @ Q.T).max() <= 1e-12 * s
    # the reference displacement enters the strain as u + uref
    ur = rng.normal(size=u.shape) * 1e-3
    ra = _res(x, ienb, z, z, u - ur, w, uref=ur)
    assert np.abs(ra - r0).max() <= 1e-12 * s


def test_mass_support_damping_closed_forms():
    x, ienb, npl, (xs, ys) = _plate(5, 3)
    p = O.make_params()
    area = np.zeros(x.shape[0])
    for f in ienb[:, :3] - 1:
        a = 0.5 * np.linalg.norm(np.cross(x[f[1]] - x[f[0]], x[f[2]] - x[f[0]]))
        area[f] += a / 3
    z = np.zeros((x.shape[0], 4))
    uz = np.zeros((x.shape[0], 3))
    acc = np.array([3.0, -2.0, 5.0])
    ac = np.zeros((x.shape[0], 4))
    ac[:, :3] = acc
    w = O.make_wall(p, iwallsupp=0, iwalldamp=0, iringsupp=0, iringdamp=0)
    r = _res(x, ienb, z, ac, uz, w)
    want = -w.rhovw * w.thicknessvw * area[:, None] * acc
    assert np.abs(r[:, :3] - want).max() <= 1e-13 * np.abs(want).max()
    # support on a rigid translation: the membrane strain vanishes
    cu = np.array([1e-2, 2e-2, -3e-2])
    w = O.make_wall(p, iwalldamp=0, iringsupp=0, iringdamp=0)
    r = _res(x, ienb, z, z, np.tile(cu, (x.shape[0], 1)), w)
    want = -w.tissSuppStiffCoeff * area[:, None] * cu
    # the membrane term of a rigid translation is rounding of E h u-sized terms
    assert np.abs(r[:, :3] - want).max() <= 1e-14 * w.evw * w.thicknessvw * np.abs(cu).max()
    # damping of a uniform velocity, and the continuity term -N_a u.n
    v = np.array([0.5, -1.5, 2.0])
    y = np.zeros((x.shape[0], 4))
    y[:, :3] = v
    w = O.make_wall(p, iwallsupp=0, iringsupp=0, iringdamp=0)
    r = _res(x, ienb, y, z, uz, w)
    want = -w.tissSuppDampCoeff * area[:, None] * v
    assert np.abs(r[:, :3] - want).max() <= 1e-13 * np.abs(want).max()
    nz = np.sign(np.cross(x[ienb[0, 1] - 1] - x[ienb[0, 0] - 1], x[ienb[0, 2] - 1] - x[ienb[0, 0] - 1])[2])
    assert np.abs(r[:npl, 3] + area[:npl] * nz * v[2]).max() <= 1e-13 * abs(v[2]) * area.max()


def test_ring_tethering_edge_integral():
    # one wall face with its first two nodes on a ring z = 0 and its third
    # node above: the tethering direction is -z (in-face, normal to the edge)
    x = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.6, 0.6, 0.9], [0.2, 0.2, 0.3]])
    ienb = np.asfortranarray([[1, 2, 3, 4]], dtype=np.int32)
    tag = np.array([1, 1, 0, 0], np.int32)
    u = np.array([[0.1, -0.2, 0.3], [0.05, 0.4, -0.7], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    z = np.zeros((4, 4))
    base = _membrane_only()
    ring = _membrane_only(iringsupp=1, iwallsupp=1)
    ring.tissSuppStiffCoeff = 0.0
    r = _res(x, ienb, z, z, u, ring, tag=tag) - _res(x, ienb, z, z, u, base, tag=tag)
    L = np.sqrt(2.0)
    d = np.array([0.0, 0.0, -1.0])
    # the in-face direction normal to the edge, towards the face's midpoint's far side
    e = (x[1] - x[0]) / L
    m = (x[0] + x[1]) / 2 - x[:3].mean(axis=0)
    d = m - e * (m @ e)
    d /= np.linalg.norm(d)
    k = ring.tissSuppRingStiffCoeff
    f1 = k * L / 6 * (2 * (u[0] @ d) + u[1] @ d) * d
    f2 = k * L / 6 * ((u[0] @ d) + 2 * (u[1] @ d)) * d
    assert np.allclose(r[0, :3], -f1, rtol=0, atol=1e-12 * np.abs(f1).max())
    assert np.allclose(r[1, :3], -f2, rtol=0, atol=1e-12 * np.abs(f1).max())
    assert np.all(r[2:] == 0.0)


def _cyl_wall():
    from crimsonflowsolver_amd import mesh as M
    m = M.cylinder_mesh(4, 4, 3)
    ienb, iBCB, BCB = M.cylinder_boundary(m)
    keep = (np.asarray(iBCB)[:, 0] & 16) != 0  # the wall faces: the others carry no tangent
    ienb = np.asfortranarray(np.asarray(ienb)[keep])
    iBCB = np.asfortranarray(np.asarray(iBCB)[keep])
    BCB = np.asfortranarray(np.asarray(BCB)[keep])
    colm, rowp = O.genadj(m.ien, m.numnp)
    tag = np.zeros(m.numnp, np.int32)
    tag[m.wall & (m.inlet | m.outlet)] = 1  # the wall's end rings
    return m, ienb, iBCB, BCB, colm, rowp, tag


@pytest.mark.parametrize("ringdamp", [0, 1])
def test_wall_tangent_is_the_residual_derivative(ringdamp):
    m, ienb, iBCB, BCB, colm, rowp, tag = _cyl_wall()
    p = O.make_params()
    # (a ring damping large enough to show beside the membrane)
    w = O.make_wall(p, iringdamp=ringdamp, tissSuppRingDampCoeff=1e4)
    N = m.numnp
    rng = np.random.default_rng(7)
    y = rng.normal(size=(N, 4))
    ac = rng.normal(size=(N, 4)) * 10
    u = rng.normal(size=(N, 3)) * 1e-3
    uref = rng.normal(size=(N, 3)) * 1e-3
    kw = dict(uref=uref, tag=tag)

    def R(t, d):
        yy, aa, uu = y.copy(), ac.copy(), u.copy()
        uu += t * p.alfi * w.betai * p.Delt ** 2 * d
        aa[:, :3] += t * p.almi * d
        yy[:, :3] += t * p.alfi * p.gami * p.Delt * d
        res = np.zeros((N, 4), order="F")
        return O.boundary_wall(ienb, iBCB, BCB, yy, aa, uu, m.x, res, w, **kw)

    K = np.zeros((9, rowp.size), order="F")
    O.boundary_wall(ienb, iBCB, BCB, y, ac, u, m.x, np.zeros((N, 4), order="F"), w, ibc=np.zeros(N, np.int32),
                    bc=np.zeros((N, 6)), colm=colm, rowp=rowp, lhsK=K, **kw)
    # the assembled 3x3 blocks as a dense (3N, 3N) matrix, K(i, j) of the
    # residual component i and the increment component j at lhsK(i + 3 j)
    rows = np.repeat(np.arange(N), np.diff(colm))
    A = np.zeros((3 * N, 3 * N))
    for c in range(3):
        for j in range(3):
            np.add.at(A, (3 * rows + c, 3 * (rowp - 1) + j), K[c + 3 * j])
    At = np.zeros((3 * N, 3 * N))
    for c in range(3):
        for j in range(3):
            np.add.at(At, (3 * rows + c, 3 * (rowp - 1) + j), K[3 * c + j])
    for trial in range(3):
        d = rng.normal(size=(N, 3))
        h = 1e-3
        fd = -(R(h, d) - R(-h, d))[:, :3] / (2 * h)
        an = (A @ d.reshape(-1)).reshape(N, 3)
        ant = (At @ d.reshape(-1)).reshape(N, 3)
        e1 = np.abs(an - fd).max() / np.abs(fd).max()
        e2 = np.abs(ant - fd).max() / np.abs(fd).max()
        print(f"ringdamp {ringdamp}: tangent read as [eq i][var j] = lhsK(i+3j) vs FD {e1:.2e}, "
              f"read transposed {e2:.2e}")
        if ringdamp == 0:
            assert e2 < 1e-8 and e1 > 1e-3
        else:
            assert e2 > 1e-3
