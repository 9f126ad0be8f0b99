"""GPU checks at the benchmark workloads' full sizes.

20x20x100 cylinder (189,600 tets, 36,057 nodes; rho 1.06, mu 0.04, Poiseuille
state with 5 % seeded noise):
  * the FP64 formation with the cylinder's boundary elements against the
    oracle over the whole mesh, normwise relative 1e-12 per array;
  * one memLS solve of that system under both readings of "GMRES dim 50,
    rel tol 1e-6": GMRESV on the 4x4 system and the NS solver, each against
    the oracle's memLS on the same system (convergence flag and iteration
    counts equal; the GPU's Gram-Schmidt is classical + DGKS where memLS's
    is modified, so the non-converging GMRESV is compared by the level its
    residual stalls at); the NS reading again with the nodes numbered at
    random (4-byte block columns in memLS's products instead of 2-byte
    offsets from the row).

64x64x320 cylinder (6,197,760 tets, 1,077,597 nodes) and the workloads built
on it, where the oracle cannot run over the whole mesh in seconds:
  * FP64 res, lhsK and lhsP rows of two slabs of nodes (mid-length and at
    the outlet, with boundary elements) against the oracle run on the
    elements within two rings of the slab, whose rows are then complete;
  * bitwise repeatability of the formation;
  * the constant-state patch test (interior rows vanish);
  * the mixed form on the same cylinder with 5 % of its elements stretched
    up to 50x against the FP64 form: residual within 1e-12; tangent within
    1e-4 with the FP64 tau island at aspect > 20 and within 1e-5 with the
    island at aspect > 2;
  * the VJP against central differences of the residual (h 1e-6 relative);
  * the 48-instance ensemble on 8 uneven cylinders: every instance's rows
    and blocks of the union formation bitwise those of the instance formed
    alone.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def env():
    import torch
    import crimsonflowsolver_amd as cfa
    if cfa.device_count() < 1:
        pytest.fail("no GPU visible: the HIP path must run here")
    return cfa, torch


def _T(torch, a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, order="F").T)).cuda()


# ------------------------------------------------------------ 20x20x100
@pytest.fixture(scope="module")
def cyl20(env, oracle_lib):
    cfa, _ = env
    from crimsonflowsolver_amd import mesh as M
    m = M.cylinder_mesh(20, 20, 100)
    y, ac = M.poiseuille_state(m)
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    ienb, iBCB, BCB = M.cylinder_boundary(m)
    ibc, bc = M.dirichlet_codes(m)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp)
    form.set_boundary(ienb, iBCB, BCB)
    res, K, P = form(y, ac, m.x)
    rres, rK, rP = oracle_lib.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, oracle_lib.make_params())
    rres = oracle_lib.boundary_res(ienb, iBCB, BCB, y, m.x, rres)
    return dict(m=m, colm=colm, rowp=rowp, ibc=ibc, bc=bc, gpu=(res, K, P), ref=(rres, rK, rP))


def test_cylinder_20x20x100_formation_matches_oracle(cyl20):
    m = cyl20["m"]
    print(f"20x20x100 cylinder: nel {m.nel}, nshg {m.numnp}, nnzb {cyl20['rowp'].size}", flush=True)
    errs = {n: _rel(a, b) for n, a, b in zip(("res", "lhsK", "lhsP"), cyl20["gpu"], cyl20["ref"])}
    print(errs)
    assert m.nel == 189600
    assert all(e <= TOL for e in errs.values()), errs


@pytest.mark.parametrize("reading", ["GMRESV", "NS"])
def test_cylinder_20x20x100_solve_matches_oracle(env, oracle_lib, cyl20, reading):
    cfa, _ = env
    m, colm, rowp, ibc = cyl20["m"], cyl20["colm"], cyl20["rowp"], cyl20["ibc"]
    rres, rK, rP = cyl20["ref"]
    R = np.asfortranarray(oracle_lib.bc3res(ibc, rres, cyl20["bc"]).T)
    val = oracle_lib.pack_val4(colm, rowp, rK, rP, mt=True)
    lt_gpu = cfa.LS_TYPE_GMRES if reading == "GMRESV" else cfa.LS_TYPE_NS
    lt_ref = oracle_lib.LS_TYPE_GMRES if reading == "GMRESV" else oracle_lib.LS_TYPE_NS
    kw = dict(relTol=1e-6, dimKry=50, maxItr=10)
    fixed = np.nonzero(ibc)[0] + 1
    zero = np.zeros((3, fixed.size), order="F")
    print(f"20x20x100 solve, {reading}", flush=True)
    ls = cfa.memLS_LS_CREATE(lt_gpu, **kw)
    lhs = cfa.memLS_LHS_CREATE(m.numnp, m.numnp, rowp.size, None, colm, rowp, 1)
    cfa.memLS_BC_CREATE(lhs, 1, fixed.size, 3, cfa.BC_TYPE_Dir, fixed, zero)
    Ri = R.copy(order="F")
    cfa.memLS_SOLVE(lhs, ls, 4, Ri, val)
    ref = oracle_lib.MemLS(colm, rowp, 1, lt_ref, mt=True, **kw)
    ref.bc_create(1, fixed, 3, oracle_lib.BC_TYPE_Dir, zero)
    xo = ref.solve(R, val)
    g, o = ls.RI, ref.stats(0)
    print("gpu", g, "oracle", o)
    assert g["suc"] == o["suc"]
    assert abs(g["iNorm"] - o["iNorm"]) <= 1e-12 * o["iNorm"]
    if reading == "NS":
        assert g["suc"] == 1
        for sub, name in ((0, "RI"), (1, "GM"), (2, "CG")):
            gi, oi = getattr(ls, name)["itr"], ref.stats(sub)["itr"]
            print(name, "itr gpu", gi, "oracle", oi)
            assert abs(gi - oi) <= max(2, 0.02 * oi), (name, gi, oi)
        assert np.abs(Ri - xo).max() <= 1e-4 * np.abs(xo).max()
    else:
        # restarted GMRES on the unsplit 4x4 system stalls on this mesh in
        # both: every restart used, and the residual left at the same level
        assert g["itr"] == o["itr"]
        ratio = (g["fNorm"] / g["iNorm"]) / (o["fNorm"] / o["iNorm"])
        print("stall level gpu/oracle", ratio)
        assert 0.5 <= ratio <= 2.0


def test_cylinder_20x20x100_ns_solve_scattered_numbering(env, oracle_lib, cyl20):
    """The NS reading on the same cylinder with its nodes numbered at random:
    some block columns lie more than 32,767 rows away, so memLS keeps 4-byte
    column indices for its products, where the numbered cylinder's products
    read 2-byte offsets from the row; against the oracle on the renumbered
    system (the same checks as the NS reading above)."""
    cfa, _ = env
    from crimsonflowsolver_amd import mesh as M
    m = cyl20["m"]
    n = m.numnp
    perm = np.random.default_rng(2020).permutation(n)  # new number of old node i
    inv = np.argsort(perm)
    ien = np.asfortranarray((perm[np.asarray(m.ien) - 1] + 1).astype(np.int32))
    x = np.asfortranarray(m.x[inv])
    y0, ac0 = M.poiseuille_state(m)
    y, ac = np.asfortranarray(y0[inv]), np.asfortranarray(ac0[inv])
    ibc = np.ascontiguousarray(cyl20["ibc"][inv])
    bc = np.asfortranarray(cyl20["bc"][inv])
    colm, rowp = cfa.genadj(ien, n)
    rows = np.repeat(np.arange(n), np.diff(colm))
    assert np.abs((rowp - 1) - rows).max() > 32767
    rres, rK, rP = oracle_lib.elmgmr_interior(ien, y, ac, x, colm, rowp, oracle_lib.make_params())
    R = np.asfortranarray(oracle_lib.bc3res(ibc, rres, bc).T)
    val = oracle_lib.pack_val4(colm, rowp, rK, rP, mt=True)
    kw = dict(relTol=1e-6, dimKry=50, maxItr=10)
    fixed = np.nonzero(ibc)[0] + 1
    zero = np.zeros((3, fixed.size), order="F")
    ls = cfa.memLS_LS_CREATE(cfa.LS_TYPE_NS, **kw)
    lhs = cfa.memLS_LHS_CREATE(n, n, rowp.size, None, colm, rowp, 1)
    cfa.memLS_BC_CREATE(lhs, 1, fixed.size, 3, cfa.BC_TYPE_Dir, fixed, zero)
    Ri = R.copy(order="F")
    cfa.memLS_SOLVE(lhs, ls, 4, Ri, val)
    ref = oracle_lib.MemLS(colm, rowp, 1, oracle_lib.LS_TYPE_NS, mt=True, **kw)
    ref.bc_create(1, fixed, 3, oracle_lib.BC_TYPE_Dir, zero)
    xo = ref.solve(R, val)
    g, o = ls.RI, ref.stats(0)
    print("scattered numbering: gpu", g, "oracle", o)
    assert g["suc"] == o["suc"] == 1
    assert abs(g["iNorm"] - o["iNorm"]) <= 1e-12 * o["iNorm"]
    for sub, name in ((0, "RI"), (1, "GM"), (2, "CG")):
        gi, oi = getattr(ls, name)["itr"], ref.stats(sub)["itr"]
        print(name, "itr gpu", gi, "oracle", oi)
        assert abs(gi - oi) <= max(2, 0.02 * oi), (name, gi, oi)
    assert np.abs(Ri - xo).max() <= 1e-4 * np.abs(xo).max()


# ------------------------------------------------------------ 64x64x320
@pytest.fixture(scope="module")
def cyl64(env):
    cfa, _ = env
    from crimsonflowsolver_amd import mesh as M
    m = M.cylinder_mesh(64, 64, 320)
    y, ac = M.poiseuille_state(m)
    rng = np.random.default_rng(64)
    ac[:, :3] = rng.uniform(-50, 50, size=(m.numnp, 3))
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    bnd = M.cylinder_boundary(m)
    print(f"64x64x320 cylinder: nel {m.nel}, nshg {m.numnp}, nnzb {rowp.size}", flush=True)
    return dict(m=m, y=y, ac=ac, colm=colm, rowp=rowp, bnd=bnd)


@pytest.fixture(scope="module")
def cyl64_formed(env, cyl64):
    cfa, _ = env
    d = cyl64
    form = cfa.ElmGMRInterior(d["m"].ien, d["colm"], d["rowp"], d["m"].numnp)
    form.set_boundary(*d["bnd"])
    return form, form(d["y"], d["ac"], d["m"].x)


def _slab_oracle(oracle_lib, m, y, ac, bnd, z0, z1):
    """The oracle's formation on the elements within two rings of the nodes
    with z in [z0, z1]; returns (slab nodes, the submesh's node list, its
    colm/rowp, res, lhsK, lhsP)."""
    ien0 = np.asarray(m.ien) - 1
    T = np.nonzero((m.x[:, 2] >= z0) & (m.x[:, 2] <= z1))[0]
    mask = np.zeros(m.numnp, bool)
    mask[T] = True
    e1 = mask[ien0].any(axis=1)
    mask1 = np.zeros(m.numnp, bool)
    mask1[ien0[e1].ravel()] = True
    e2 = mask1[ien0].any(axis=1)
    nodes = np.unique(ien0[e2].ravel())
    sub = np.full(m.numnp, -1, np.int64)
    sub[nodes] = np.arange(nodes.size)
    sien = np.asfortranarray((sub[ien0[e2]] + 1).astype(np.int32))
    sx, sy, sac = (np.asfortranarray(a[nodes]) for a in (m.x, y, ac))
    colm, rowp = oracle_lib.genadj(sien, nodes.size)
    res, K, P = oracle_lib.elmgmr_interior(sien, sy, sac, sx, colm, rowp, oracle_lib.make_params())
    ienb, iBCB, BCB = bnd
    inb = (sub[np.asarray(ienb) - 1] >= 0).all(axis=1)
    if inb.any():
        sb = np.asfortranarray((sub[np.asarray(ienb)[inb] - 1] + 1).astype(np.int32))
        res = oracle_lib.boundary_res(sb, np.asfortranarray(iBCB[inb]), np.asfortranarray(BCB[inb]), sy, sx, res)
    return T, sub, colm, rowp, res, K, P, int(inb.sum())


def _row_ranges(colm, rows):
    lo, hi = colm[rows] - 1, colm[rows + 1] - 1
    return np.concatenate([np.arange(a, b) for a, b in zip(lo, hi)])


@pytest.mark.parametrize("z0,z1", [(5.0, 5.07), (9.93, 10.0)])
def test_cylinder_64x64x320_slab_rows_match_oracle(oracle_lib, cyl64, cyl64_formed, z0, z1):
    d = cyl64
    m, colm, rowp = d["m"], d["colm"], d["rowp"]
    _, (res, K, P) = cyl64_formed
    T, sub, scolm, srowp, rres, rK, rP, nb = _slab_oracle(oracle_lib, m, d["y"], d["ac"], d["bnd"], z0, z1)
    print(f"slab z in [{z0}, {z1}]: {T.size} rows, {nb} boundary elements in the oracle's submesh", flush=True)
    gi = _row_ranges(colm, T)
    si = _row_ranges(scolm, sub[T])
    assert gi.size == si.size
    assert np.array_equal(rowp[gi] - 1, np.flatnonzero(sub >= 0)[srowp[si] - 1])
    errs = {"res": _rel(res[T], rres[sub[T]]), "lhsK": _rel(K[:, gi], rK[:, si]), "lhsP": _rel(P[:, gi], rP[:, si])}
    print(errs)
    assert all(e <= TOL for e in errs.values()), errs


def test_cylinder_64x64x320_formation_repeatable_bitwise(env, cyl64, cyl64_formed):
    cfa, torch = env
    d = cyl64
    form, _ = cyl64_formed
    m = d["m"]
    ty, tac, tx = _T(torch, d["y"]), _T(torch, d["ac"]), _T(torch, m.x)
    outs = []
    for _ in range(2):
        r = torch.empty((4, m.numnp), dtype=torch.float64, device="cuda")
        K = torch.empty((d["rowp"].size, 9), dtype=torch.float64, device="cuda")
        P = torch.empty((d["rowp"].size, 4), dtype=torch.float64, device="cuda")
        form.dev(ty, tac, tx, r, K, P)
        torch.cuda.synchronize()
        outs.append((r, K, P))
    for a, b in zip(*outs):
        assert torch.equal(a, b)


def test_cylinder_64x64x320_patch_test(env, cyl64):
    """Constant velocity and pressure, zero acceleration: every element's
    residual sums to zero over a closed patch, so interior rows vanish to
    rounding (the boundary rows keep their O(V p |grad N|) values)."""
    cfa, torch = env
    d = cyl64
    m = d["m"]
    y = np.zeros((m.numnp, 4), order="F")
    y[:] = [1.5, -0.7, 20.0, 1000.0]
    ac = np.zeros((m.numnp, 4), order="F")
    form = cfa.ElmGMRInterior(m.ien, d["colm"], d["rowp"], m.numnp)
    res, _, _ = form(y, ac, m.x, lhs=0)
    interior = ~(m.inlet | m.outlet | m.wall)
    ratio = np.abs(res[interior]).max() / np.abs(res[~interior]).max()
    print("interior / boundary max |res|", ratio)
    assert ratio <= 1e-11


@pytest.fixture(scope="module")
def stretched64(cyl64):
    from crimsonflowsolver_amd import mesh as M
    return M.stretch_regions(cyl64["m"])


def test_cylinder_64x64x320_stretched_mixed_against_fp64(env, cyl64, stretched64):
    """The mixed form against FP64 on the stretched cylinder.  Elements
    below the island's aspect threshold are formed in FP32, whose metric
    loses about aspect^2 ulps: at the workload's threshold 20 the tangent's
    error is stated and bounded by 1e-4; with the threshold at 2 (every
    stretched element in the FP64 island) it falls by a factor of about 10,
    to the FP32 level (bounded by 1e-5 here: the maximum over 6.2 M elements
    with accelerations of +/-50).  The residual stays FP64 (1e-12)."""
    cfa, torch = env
    from crimsonflowsolver_amd import mesh as M
    d = cyl64
    ms = stretched64
    moved = M.moved_elements(d["m"], ms)
    asp = M.aspect_ratios(ms.x, np.asarray(ms.ien)[moved])
    frac = moved.size / ms.nel
    print(f"stretched: {frac:.3f} of the elements moved, max aspect {asp.max():.1f}, "
          f"{(asp > 20).sum()} above 20", flush=True)
    assert 0.04 <= frac and asp.max() > 40.0
    ty, tac, tx = _T(torch, d["y"]), _T(torch, d["ac"]), _T(torch, ms.x)

    def formed(prec, thr):
        f = cfa.ElmGMRInterior(ms.ien, d["colm"], d["rowp"], ms.numnp,
                               params=cfa.form_params(precision=prec, aspect_fp64=thr))
        r = torch.empty((4, ms.numnp), dtype=torch.float64, device="cuda")
        K = torch.empty((d["rowp"].size, 9), dtype=torch.float64, device="cuda")
        P = torch.empty((d["rowp"].size, 4), dtype=torch.float64, device="cuda")
        f.dev(ty, tac, tx, r, K, P)
        torch.cuda.synchronize()
        return r, K, P

    def rel(a, b):
        return float((a - b).abs().max() / b.abs().max())

    ref = formed(64, 20.0)
    for thr, bar in ((20.0, 1e-4), (2.0, 1e-5)):
        out = formed(32, thr)
        errs = {n: rel(b, a) for n, a, b in zip(("res", "lhsK", "lhsP"), ref, out)}
        print(f"island threshold {thr}: {errs}", flush=True)
        assert errs["res"] <= TOL
        assert errs["lhsK"] <= bar and errs["lhsP"] <= bar
        assert errs["lhsK"] > 1e-10  # the tangent really was formed in FP32
        del out


def test_cylinder_64x64x320_vjp_central_differences(env, cyl64):
    cfa, torch = env
    d = cyl64
    m = d["m"]
    form = cfa.ElmGMRInterior(m.ien, d["colm"], d["rowp"], m.numnp)
    form.set_boundary(*d["bnd"])
    rng = np.random.default_rng(77)
    ty, tac, tx = _T(torch, d["y"]), _T(torch, d["ac"]), _T(torch, m.x)
    w = _T(torch, rng.normal(size=(m.numnp, 4)))
    ybar = torch.empty_like(ty)
    mub = form.vjp(ty, tac, tx, w, ybar)
    r = torch.empty_like(ty)

    def J(yy, mu=None):
        # another viscosity: every element (and boundary element) as one
        # ensemble instance of that viscosity
        if mu is not None:
            form.set_instances(np.zeros(m.nel, np.uint16), [mu])
        form.dev(yy, tac, tx, r, lhs=0)
        torch.cuda.synchronize()
        if mu is not None:
            form.set_instances(None, [])
        return float((w * r).sum())

    dirn = _T(torch, rng.normal(size=(m.numnp, 4)) * np.array([1.0, 1.0, 1.0, 30.0]))
    h = 1e-6 * float(ty.abs().max()) / float(dirn.abs().max())
    fd = (J(ty + h * dirn) - J(ty - h * dirn)) / (2 * h)
    an = float((ybar * dirn).sum())
    print("state: fd", fd, "vjp", an)
    assert abs(fd - an) <= 1e-5 * abs(fd)
    hm = 1e-6 * 0.04
    fdm = (J(ty, 0.04 + hm) - J(ty, 0.04 - hm)) / (2 * hm)
    print("mu: fd", fdm, "vjp", mub, flush=True)
    assert abs(fdm - mub) <= 1e-5 * abs(fdm)


def test_ensemble_48_union_equals_instances(env):
    cfa, torch = env
    from crimsonflowsolver_amd import mesh as M
    from crimsonflowsolver_amd.ensemble import Ensemble, uneven_cylinders
    meshes, shapes, mu, R = uneven_cylinders(48)
    ens = Ensemble(meshes, mu, R, boundary=True)
    print(f"ensemble: 48 instances, {ens.elem_off[-1]} tets, {ens.N} nodes, {ens.nelb} boundary elements",
          flush=True)
    st = [M.poiseuille_state(mm, seed=60 + k) for k, mm in enumerate(meshes)]
    ty = _T(torch, np.concatenate([a[0] for a in st]))
    tac = _T(torch, np.concatenate([a[1] for a in st]))
    tx = _T(torch, ens.x)
    nnz = ens.rowp.size
    ru = torch.empty((4, ens.N), dtype=torch.float64, device="cuda")
    KU = torch.empty((nnz, 9), dtype=torch.float64, device="cuda")
    PU = torch.empty((nnz, 4), dtype=torch.float64, device="cuda")
    ens.form.dev(ty, tac, tx, ru, KU, PU)
    torch.cuda.synchronize()
    handles = {}
    for k, mm in enumerate(meshes):
        if id(mm) not in handles:
            colm, rowp = cfa.genadj(mm.ien, mm.numnp)
            f = cfa.ElmGMRInterior(mm.ien, colm, rowp, mm.numnp)
            f.set_boundary(*M.cylinder_boundary(mm))
            handles[id(mm)] = (f, rowp.size)
        f, nk = handles[id(mm)]
        f.set_instances(np.zeros(mm.nel, np.uint16), mu[k:k + 1])
        r = torch.empty((4, mm.numnp), dtype=torch.float64, device="cuda")
        K = torch.empty((nk, 9), dtype=torch.float64, device="cuda")
        P = torch.empty((nk, 4), dtype=torch.float64, device="cuda")
        f.dev(_T(torch, st[k][0]), _T(torch, st[k][1]), _T(torch, mm.x), r, K, P)
        torch.cuda.synchronize()
        rows, blk = ens.rows(k), ens.nnz_range(k)
        assert torch.equal(ru[:, rows], r), k
        assert torch.equal(KU[blk], K) and torch.equal(PU[blk], P), k
