"""GPU parity of the formation path (crimson_form_* through the C ABI) against
the oracle (the CPU restatement of elmgmr.f90/e3*.f90/fillsparse.f90).

Tolerance: FP64, normwise relative per output array, max|gpu - oracle| /
max|oracle| <= 1e-12 (the GPU sums the same terms in a different order).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def cfa():
    import crimsonflowsolver_amd as cfa
    if cfa.device_count() < 1:
        pytest.fail("no GPU visible: the HIP path must run here")
    return cfa


def _case(grid, seed, noise=0.05):
    from crimsonflowsolver_amd import mesh as M
    m = M.cylinder_mesh(*grid)
    y, ac = M.poiseuille_state(m, seed=seed, noise=noise)
    rng = np.random.default_rng(seed)
    ac[:, :3] = rng.uniform(-50, 50, size=(m.numnp, 3))  # nonzero acceleration
    return m, y, ac


@pytest.mark.parametrize("grid,seed", [((4, 4, 6), 1), ((6, 6, 10), 2), ((10, 10, 30), 3)])
def test_formation_matches_oracle(cfa, oracle_lib, grid, seed):
    m, y, ac = _case(grid, seed)
    colm, rowp = oracle_lib.genadj(m.ien, m.numnp)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp)
    res, K, P, q = form(y, ac, m.x, want_q=True)
    rres, rK, rP, rq = oracle_lib.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, oracle_lib.make_params(),
                                                  want_q=True)
    errs = {n: _rel(a, b) for n, a, b in (("res", res, rres), ("lhsK", K, rK), ("lhsP", P, rP), ("q", q, rq))}
    print(grid, seed, errs)
    assert all(e <= TOL for e in errs.values()), errs


@pytest.mark.parametrize("nel", [1, 2])
def test_formation_smallest_meshes(cfa, oracle_lib, nel):
    """One tet (4 rows, one incidence each) and two tets sharing a face
    (rows of one and two incidences): the smallest inputs, every grid one
    workgroup."""
    rng = np.random.default_rng(40 + nel)
    x = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.2, 1.0, 0.1], [0.1, 0.2, 1.0], [0.9, 0.9, 0.8]])
    x += rng.uniform(-0.05, 0.05, x.shape)
    ien = np.array([[1, 2, 3, 4], [2, 3, 4, 5]], np.int32)[:nel]
    # positive orientation: det[x1-x4, x2-x4, x3-x4] > 0 (p1 volume sign)
    for e in range(nel):
        v = x[ien[e, :3] - 1] - x[ien[e, 3] - 1]
        if np.linalg.det(v) < 0:
            ien[e, [0, 1]] = ien[e, [1, 0]]
    ien = np.asfortranarray(ien)
    n = 4 + (nel - 1)
    x = np.asfortranarray(x[:n])
    y = np.asfortranarray(rng.normal(size=(n, 4)) * np.array([10.0, 1.0, 1.0, 1.0]))
    ac = np.asfortranarray(rng.uniform(-20, 20, size=(n, 4)))
    colm, rowp = oracle_lib.genadj(ien, n)
    form = cfa.ElmGMRInterior(ien, colm, rowp, n)
    res, K, P, q = form(y, ac, x, want_q=True)
    rres, rK, rP, rq = oracle_lib.elmgmr_interior(ien, y, ac, x, colm, rowp, oracle_lib.make_params(),
                                                  want_q=True)
    errs = {n: _rel(a, b) for n, a, b in (("res", res, rres), ("lhsK", K, rK), ("lhsP", P, rP), ("q", q, rq))}
    print("nel", nel, errs)
    assert all(e <= TOL for e in errs.values()), errs


def test_formation_residual_only(cfa, oracle_lib):
    m, y, ac = _case((6, 6, 8), 7)
    colm, rowp = oracle_lib.genadj(m.ien, m.numnp)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp)
    res, K, P = form(y, ac, m.x, lhs=0)
    assert K is None and P is None
    rres, _, _ = oracle_lib.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, oracle_lib.make_params(lhs=0))
    assert _rel(res, rres) <= TOL


def test_formation_repeatable_bitwise(cfa, oracle_lib):
    """Deterministic assembly: two calls give identical bits."""
    m, y, ac = _case((8, 8, 12), 11)
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp)
    a = form(y, ac, m.x)
    b = form(y, ac, m.x)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_fused_val4_matches_solfar_packing(cfa, oracle_lib):
    """Formation straight into the memLS matrix equals solfar's packing of
    the oracle's lhsK/lhsP (solfar.f90:212-235)."""
    import torch
    m, y, ac = _case((6, 6, 10), 5)
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp)
    lhs = cfa.memLS_LHS_CREATE(m.numnp, m.numnp, rowp.size, None, colm, rowp, 1)
    dev = torch.device("cuda:0")
    ty = torch.from_numpy(np.ascontiguousarray(y.T)).to(dev)    # (4,nshg) C == (nshg,4) F
    tac = torch.from_numpy(np.ascontiguousarray(ac.T)).to(dev)
    tx = torch.from_numpy(np.ascontiguousarray(m.x.T)).to(dev)
    tres4 = torch.zeros((m.numnp, 4), dtype=torch.float64, device=dev)  # Res4(4,nshg) Fortran
    form.to_lhs(ty, tac, tx, tres4, lhs)
    torch.cuda.synchronize()
    val = lhs.get_val(4)
    rres, rK, rP = oracle_lib.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, oracle_lib.make_params())
    rval = oracle_lib.pack_val4(colm, rowp, rK, rP)
    assert _rel(val, rval) <= TOL
    assert _rel(tres4.cpu().numpy(), rres) <= TOL


def test_device_entry_point(cfa, oracle_lib):
    import torch
    m, y, ac = _case((5, 5, 7), 9)
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp)
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).to(dev)
    tres = torch.zeros((4, m.numnp), dtype=torch.float64, device=dev)
    tK = torch.zeros((rowp.size, 9), dtype=torch.float64, device=dev)
    tP = torch.zeros((rowp.size, 4), dtype=torch.float64, device=dev)
    form.dev(T(y), T(ac), T(m.x), tres, tK, tP)
    rres, rK, rP = oracle_lib.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, oracle_lib.make_params())
    assert _rel(tres.cpu().numpy().T, rres) <= TOL
    assert _rel(tK.cpu().numpy().T, rK) <= TOL
    assert _rel(tP.cpu().numpy().T, rP) <= TOL


def test_bc3res_device(cfa, oracle_lib):
    import torch
    from crimsonflowsolver_amd import mesh as M
    m, y, ac = _case((5, 5, 6), 13)
    ibc, bc = M.dirichlet_codes(m)
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, iBC=ibc)
    res, _, _ = form(y, ac, m.x, lhs=0)
    t = torch.from_numpy(np.ascontiguousarray(res.T)).cuda()
    form.bc3res(t)
    want = oracle_lib.bc3res(ibc, res, bc)
    assert np.array_equal(t.cpu().numpy().T, want)


def test_formation_unstructured_matches_oracle(cfa, oracle_lib):
    """A jittered Delaunay mesh: nodes with up to ~50 incident elements and
    ~32 neighbours, so rows take several incidence chunks and several
    16-block groups in the assembly (the Kuhn meshes have 24 and 15).  Host
    layouts, the fused memLS matrix, and the FP32 tangent."""
    import torch
    from crimsonflowsolver_amd import mesh as M
    m = M.jittered_delaunay_mesh(7, 7, 12, seed=17)
    y, ac = M.poiseuille_state(m, seed=17, noise=0.2)
    rng = np.random.default_rng(17)
    ac[:, :3] = rng.uniform(-50, 50, size=(m.numnp, 3))
    colm, rowp = oracle_lib.genadj(m.ien, m.numnp)
    assert np.diff(colm).max() > 16 and np.bincount(m.ien.ravel()).max() > 32
    rres, rK, rP, rq = oracle_lib.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, oracle_lib.make_params(),
                                                  want_q=True)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp)
    res, K, P, q = form(y, ac, m.x, want_q=True)
    errs = {n: _rel(a, b) for n, a, b in (("res", res, rres), ("lhsK", K, rK), ("lhsP", P, rP), ("q", q, rq))}
    print("host", errs)
    assert all(e <= TOL for e in errs.values()), errs
    lhs = cfa.memLS_LHS_CREATE(m.numnp, m.numnp, rowp.size, None, colm, rowp, 1)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
    tres4 = torch.zeros((m.numnp, 4), dtype=torch.float64, device="cuda")
    form.to_lhs(T(y), T(ac), T(m.x), tres4, lhs)
    rval = oracle_lib.pack_val4(colm, rowp, rK, rP)
    print("fused", _rel(lhs.get_val(4), rval), _rel(tres4.cpu().numpy(), rres))
    assert _rel(lhs.get_val(4), rval) <= TOL
    assert _rel(tres4.cpu().numpy(), rres) <= TOL
    form32 = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, params=cfa.form_params(precision=32))
    res32, K32, P32 = form32(y, ac, m.x)
    print("fp32 tangent", _rel(K32, rK), _rel(P32, rP))
    assert _rel(res32, rres) <= TOL
    assert _rel(K32, rK) <= 3e-6 and _rel(P32, rP) <= 3e-6


def test_formation_real_geometry(cfa, oracle_lib):
    """The reference's own test cylinder (mainTests/basic: 1072 nodes, 4157
    tets, tests/golden/basic_case.npz) with a seeded Poiseuille state along
    its axis: residual, tangent and q against the oracle."""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "basic_case.npz"))
    x, ien = np.asfortranarray(z["x"]), np.asfortranarray(z["ien"])
    n = x.shape[0]
    rng = np.random.default_rng(23)
    r2 = (x[:, 0] ** 2 + x[:, 1] ** 2) / 9.0
    y = np.zeros((n, 4), order="F")
    y[:, :3] = rng.uniform(-1.5, 1.5, size=(n, 3))
    y[:, 2] += 30.0 * np.clip(1.0 - r2, 0.0, None)
    y[:, 3] = 1330.0 * (20.0 - x[:, 2]) / 40.0
    ac = np.asfortranarray(rng.uniform(-50, 50, size=(n, 4)))
    colm, rowp = oracle_lib.genadj(ien, n)
    rres, rK, rP, rq = oracle_lib.elmgmr_interior(ien, y, ac, x, colm, rowp, oracle_lib.make_params(), want_q=True)
    form = cfa.ElmGMRInterior(ien, colm, rowp, n)
    res, K, P, q = form(y, ac, x, want_q=True)
    errs = {k: _rel(a, b) for k, a, b in (("res", res, rres), ("lhsK", K, rK), ("lhsP", P, rP), ("q", q, rq))}
    print("real geometry", errs)
    assert all(e <= TOL for e in errs.values()), errs


def test_formation_scattered_numbering(cfa, oracle_lib):
    """A cylinder whose node numbers are randomly permuted (elements span
    far more than 32K numbers, so the assembly schedules rows in Morton
    order of the coordinates instead of the numbering): residual, tangent
    and the memLS planes against the oracle."""
    import torch
    from crimsonflowsolver_amd import mesh as M
    m = M.cylinder_mesh(32, 32, 80)
    n = m.numnp
    perm = np.random.default_rng(77).permutation(n)  # new number of old node i
    inv = np.argsort(perm)
    ien = np.asfortranarray((perm[m.ien - 1] + 1).astype(np.int32))
    x = np.asfortranarray(m.x[inv])
    y0, ac0 = M.poiseuille_state(m, seed=78)
    y, ac = np.asfortranarray(y0[inv]), np.asfortranarray(ac0[inv])
    span = ien.max(axis=1) - ien.min(axis=1)
    assert (span > 32768).mean() > 0.01  # the numbering is not local
    colm, rowp = cfa.genadj(ien, n)
    form = cfa.ElmGMRInterior(ien, colm, rowp, n)
    res, K, P = form(y, ac, x)
    rres, rK, rP = oracle_lib.elmgmr_interior(ien, y, ac, x, colm, rowp, oracle_lib.make_params())
    errs = {"res": _rel(res, rres), "lhsK": _rel(K, rK), "lhsP": _rel(P, rP)}
    print("scattered numbering", n, "nodes", errs)
    assert all(e <= TOL for e in errs.values()), errs
    lhs = cfa.memLS_LHS_CREATE(n, n, rowp.size, None, colm, rowp, 1)
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).to(dev)
    tres4 = torch.zeros((n, 4), dtype=torch.float64, device=dev)
    form.to_lhs(T(y), T(ac), T(x), tres4, lhs)
    torch.cuda.synchronize()
    assert _rel(lhs.get_val(4), oracle_lib.pack_val4(colm, rowp, rK, rP)) <= TOL
    assert _rel(tres4.cpu().numpy(), rres) <= TOL


@pytest.mark.parametrize("precision", [64, 32])
def test_autotune_candidates_bitwise(cfa, oracle_lib, precision):
    """Every launch configuration the autotuner may pick (row schedule:
    numbering, Morton, Morton-ordered runs; assembly chunk; XCD dealing) forms bitwise the same residual and memLS
    matrix, on a jittered Delaunay mesh whose rows take 1-3 chunks of 16 and
    1-6 of 8 (ragged last chunks), and on a cylinder of 1,100+ assembly
    workgroups (so that dealing in turns of 64 remaps all but the last
    partial turn); the FP64 ones match the oracle.  A full run returns a
    candidate and a time for each one tried."""
    import torch
    from crimsonflowsolver_amd import mesh as M
    m = M.jittered_delaunay_mesh(7, 7, 12, seed=19)
    y, ac = M.poiseuille_state(m, seed=19, noise=0.2)
    rng = np.random.default_rng(19)
    ac[:, :3] = rng.uniform(-50, 50, size=(m.numnp, 3))
    nin = np.bincount(m.ien.ravel())[1:]
    assert nin.max() > 32 and (nin % 8).any() and (nin % 16).any()
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, params=cfa.form_params(precision=precision))
    lhs = cfa.memLS_LHS_CREATE(m.numnp, m.numnp, rowp.size, None, colm, rowp, 1)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
    ty, tac, tx = T(y), T(ac), T(m.x)
    tres4 = torch.zeros((m.numnp, 4), dtype=torch.float64, device="cuda")
    names = [n for i, n in enumerate(cfa.ElmGMRInterior.AUTOTUNE_CANDIDATES) if precision == 64 or not (i // 3) % 2]
    outs = []
    for name in names:
        chosen, times = form.autotune(ty, tac, tx, tres4, lhs, reps=1, only=[name])
        assert chosen == name and set(times) == {name}
        tres4.zero_()
        form.to_lhs(ty, tac, tx, tres4, lhs)
        torch.cuda.synchronize()
        outs.append((name, tres4.cpu().numpy(), lhs.get_val(4)))
    for name, r, v in outs[1:]:
        assert np.array_equal(r, outs[0][1]), name
        assert np.array_equal(v, outs[0][2]), name
    rres, rK, rP = oracle_lib.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, oracle_lib.make_params())
    assert _rel(outs[0][1], rres) <= TOL
    if precision == 64:
        assert _rel(outs[0][2], oracle_lib.pack_val4(colm, rowp, rK, rP)) <= TOL
    mc = M.cylinder_mesh(16, 16, 80)
    assert mc.numnp * 16 // 256 > 8 * 64 * 2 and (mc.numnp * 16 // 256) % (8 * 64)
    yc, acc = M.poiseuille_state(mc, seed=20)
    cc, rc = cfa.genadj(mc.ien, mc.numnp)
    fc = cfa.ElmGMRInterior(mc.ien, cc, rc, mc.numnp, params=cfa.form_params(precision=precision))
    lc = cfa.memLS_LHS_CREATE(mc.numnp, mc.numnp, rc.size, None, cc, rc, 1)
    tc = [T(a) for a in (yc, acc, mc.x)]
    rc4 = torch.zeros((mc.numnp, 4), dtype=torch.float64, device="cuda")
    cyl = []
    for name in names:
        fc.autotune(*tc, rc4, lc, reps=1, only=[name])
        rc4.zero_()
        fc.to_lhs(*tc, rc4, lc)
        torch.cuda.synchronize()
        cyl.append((name, rc4.cpu().numpy(), lc.get_val(4)))
    for name, r, v in cyl[1:]:
        assert np.array_equal(r, cyl[0][1]), name
        assert np.array_equal(v, cyl[0][2]), name
    chosen, times = form.autotune(ty, tac, tx, tres4, lhs, reps=2)
    print("autotune", precision, chosen, times)
    assert chosen in names and set(times) == set(names) and all(t > 0 for t in times.values())
    if precision == 32:  # the FP32 assembly has no chunk choice
        with pytest.raises(cfa.CrimsonError):
            form.autotune(ty, tac, tx, tres4, lhs, only=["natural/8/range"])
