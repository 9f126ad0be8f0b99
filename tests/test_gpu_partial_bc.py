"""GPU parity of the partial velocity BCs (iBC codes 1-6: bc3lhs on the
tangent, bc3lhs.f90:41-690, applied by ElmGMR at elmgmr.f90:223-226 and
:364-372; bc3Res's projections, bc3res.f90:34-80) against the oracle,
FP64, normwise relative 1e-12; FP32 tangent 3e-6.
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


def _case(seed, grid=(5, 5, 8)):
    """Codes 1-5 on wall nodes, code 6 on nodes sharing no element with
    another partial node, pressure bit on a few, 3-component inlet."""
    from crimsonflowsolver_amd import mesh as M
    m = M.cylinder_mesh(*grid)
    rng = np.random.default_rng(seed)
    codes = np.zeros(m.numnp, np.int32)
    wall = np.nonzero(m.wall & ~m.inlet & ~m.outlet)[0]
    pick = rng.choice(wall, size=wall.size // 2, replace=False)
    codes[pick] = rng.choice([1, 2, 3, 4, 5], pick.size)
    ien0 = m.ien - 1
    for cand in rng.permutation(m.numnp):
        if codes[cand]:
            continue
        elems = np.nonzero((ien0 == cand).any(axis=1))[0]
        if not (codes[ien0[elems]] != 0).any():
            codes[cand] = 6
            if (codes == 6).sum() >= 4:
                break
    ibc = (codes << 3).astype(np.int32)
    ibc[m.inlet] = 56
    ibc[rng.choice(m.numnp, 10, replace=False)] |= 4
    bc = np.asfortranarray(rng.uniform(-1.5, 1.5, (m.numnp, 7)))
    y, ac = M.poiseuille_state(m, seed=seed)
    ac[:, :3] = rng.uniform(-50, 50, size=(m.numnp, 3))
    return m, ibc, bc, y, ac


def test_partial_bc_tangent_matches_oracle(cfa, oracle_lib):
    import torch
    m, ibc, bc, y, ac = _case(11)
    assert (ibc >> 3 == 6).sum() >= 1 and np.isin(ibc >> 3, [1, 2, 3, 4, 5]).sum() >= 20
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, iBC=ibc)
    with pytest.raises(cfa.CrimsonError) as e:  # the tangent needs BC first
        form(y, ac, m.x)
    assert e.value.code == 8
    res0, _, _ = form(y, ac, m.x, lhs=0)  # the residual does not
    form.set_bc(bc)
    res, K, P = form(y, ac, m.x)
    rres, rK, rP = oracle_lib.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, oracle_lib.make_params(),
                                              ibc=ibc, bc=bc)
    errs = {"res": _rel(res, rres), "lhsK": _rel(K, rK), "lhsP": _rel(P, rP)}
    print("partial BCs", errs)
    assert all(v <= TOL for v in errs.values()), errs
    assert _rel(res0, rres) <= TOL  # the residual-only path (other summation order)
    # straight into the memLS matrix (Val4 planes)
    lhs = cfa.memLS_LHS_CREATE(m.numnp, m.numnp, rowp.size, None, colm, rowp, 1)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).cuda()
    r4 = torch.zeros((m.numnp, 4), dtype=torch.float64, device="cuda")
    form.to_lhs(T(y), T(ac), T(m.x), r4, lhs)
    torch.cuda.synchronize()
    assert _rel(lhs.get_val(4), oracle_lib.pack_val4(colm, rowp, rK, rP)) <= TOL
    # FP32 tangent
    f32 = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, iBC=ibc, params=cfa.form_params(precision=32))
    f32.set_bc(bc)
    _, K32, P32 = f32(y, ac, m.x)
    assert _rel(K32, rK) <= 3e-6 and _rel(P32, rP) <= 3e-6
    # BC removed: the tangent is refused again
    form.set_bc(None)
    with pytest.raises(cfa.CrimsonError):
        form(y, ac, m.x)


def test_partial_bc_with_boundary_elements(cfa, oracle_lib):
    """The boundary loop's bc3lhs adds one per boundary element at the
    constrained diagonal entries of its partial nodes."""
    from crimsonflowsolver_amd import mesh as M
    m, ibc, bc, y, ac = _case(12)
    ienb, iBCB, BCB = M.cylinder_boundary(m)
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, iBC=ibc)
    form.set_bc(bc)
    form.set_boundary(ienb, iBCB, BCB)
    res, K, P = form(y, ac, m.x)
    rres, rK, rP = oracle_lib.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, oracle_lib.make_params(),
                                              ibc=ibc, bc=bc)
    rK = oracle_lib.boundary_lhs(ienb, ibc, bc, colm, rowp, rK)
    rres = oracle_lib.boundary_res(ienb, iBCB, BCB, y, m.x, rres)
    errs = {"res": _rel(res, rres), "lhsK": _rel(K, rK), "lhsP": _rel(P, rP)}
    print("partial BCs + boundary", errs)
    assert all(v <= TOL for v in errs.values()), errs
    # order of the calls does not matter: boundary first, then BC
    f2 = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, iBC=ibc)
    f2.set_boundary(ienb, iBCB, BCB)
    f2.set_bc(bc)
    _, K2, _ = f2(y, ac, m.x)
    assert np.array_equal(K2, K)


def test_bc3res_partial_codes_device(cfa, oracle_lib):
    import torch
    m, ibc, bc, y, ac = _case(13)
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, iBC=ibc)
    res = np.asfortranarray(np.random.default_rng(13).standard_normal((m.numnp, 4)))
    t = torch.from_numpy(np.ascontiguousarray(res.T)).cuda()
    with pytest.raises(cfa.CrimsonError):
        form.bc3res(t)
    form.set_bc(bc)
    form.bc3res(t)
    want = oracle_lib.bc3res(ibc, res, bc)
    assert _rel(t.cpu().numpy().T, want) <= 1e-15
    t4 = torch.from_numpy(np.ascontiguousarray(res)).cuda()  # Res4(4,nshg)
    form.bc3res(t4, res4_layout=True)
    assert _rel(t4.cpu().numpy(), want) <= 1e-15


def test_itr_bc_partial_codes_device(cfa, oracle_lib):
    """itrBC on the device for codes 1-7 and the pressure bit (itrbc.f90)."""
    import torch
    m, ibc, bc, y, ac = _case(14)
    codes = np.random.default_rng(14).integers(0, 8, m.numnp)
    ibc = ((codes << 3) | (ibc & 4)).astype(np.int32)  # every code, no element restriction needed here
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).cuda()
    ty = T(y)
    form.itr_bc(torch.from_numpy(ibc).cuda(), T(bc), ty)
    form.sync()
    want = oracle_lib.itr_bc(ibc, bc, y)
    assert _rel(ty.cpu().numpy().T, want) <= 1e-15
