"""Ensembles on the GPU: instances of uneven size and viscosity formed as one
union mesh (crimson_form_set_instances), captured in a HIP graph and solved
by one NS solve.

Checks: each instance's matrix blocks and residual rows in the union are
bit-for-bit those of the instance formed alone; a graph replay equals the
direct formation bit for bit; the union solve agrees with the instances'
own solves to the solver tolerance (1e-4 relative at NS relTol 1e-6: the
union's Krylov space spans all instances, so iterates differ).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import crimsonflowsolver_amd as cfa
    if cfa.device_count() < 1:
        pytest.fail("no GPU visible: the HIP path must run here")
    return cfa, torch


def _instances():
    from crimsonflowsolver_amd import mesh as M
    meshes = [M.cylinder_mesh(5, 5, 6), M.cylinder_mesh(6, 6, 9), M.cylinder_mesh(4, 4, 12)]
    mu = np.array([0.03, 0.04, 0.06])
    R = np.array([800.0, 1333.0, 2000.0])
    states = [M.poiseuille_state(m, seed=50 + k) for k, m in enumerate(meshes)]
    return meshes, mu, R, states


def _T(torch, a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, order="F").T)).cuda()


@pytest.mark.parametrize("precision,boundary", [(64, False), (32, False), (64, True), (32, True)])
def test_union_formation_equals_instances(env, precision, boundary):
    """Each instance's rows and blocks of the union formation are bitwise
    those of the instance formed alone (with its boundary elements, in its
    own viscosity, when boundary)."""
    cfa, torch = env
    from crimsonflowsolver_amd.ensemble import Ensemble
    from crimsonflowsolver_amd import mesh as M
    meshes, mu, R, states = _instances()
    ens = Ensemble(meshes, mu, R, precision=precision, boundary=boundary)
    y = np.concatenate([s[0] for s in states])
    ac = np.concatenate([s[1] for s in states])
    ty, tac, tx = _T(torch, y), _T(torch, ac), _T(torch, ens.x)
    r4 = torch.zeros((ens.N, 4), dtype=torch.float64, device="cuda")
    ens.form.to_lhs(ty, tac, tx, r4, ens.lhs)
    val, res = ens.lhs.get_val(4), r4.cpu().numpy()
    for k, m in enumerate(meshes):
        colm, rowp = cfa.genadj(m.ien, m.numnp)
        f = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, params=cfa.form_params(mu=float(mu[k]),
                                                                                   precision=precision))
        if boundary:
            f.set_boundary(*M.cylinder_boundary(m))
        lhs = cfa.memLS_LHS_CREATE(m.numnp, m.numnp, rowp.size, None, colm, rowp, 0)
        rk = torch.zeros((m.numnp, 4), dtype=torch.float64, device="cuda")
        f.to_lhs(_T(torch, states[k][0]), _T(torch, states[k][1]), _T(torch, m.x), rk, lhs)
        assert np.array_equal(val[:, ens.nnz_range(k)], lhs.get_val(4)), k
        assert np.array_equal(res[ens.rows(k)], rk.cpu().numpy()), k
    # the captured graph replays the same formation
    g = ens.capture(ty, tac, tx, r4)
    r4.zero_()
    g.launch()
    assert np.array_equal(r4.cpu().numpy(), res)
    assert np.array_equal(ens.lhs.get_val(4), val)


def test_graph_refuses_replay_after_its_buffers_change(env):
    """New viscosities with the same instance count keep the captured graph
    valid (a replay equals the direct formation with them); changing the
    instance count, or dropping the instances, makes crimson_graph_launch
    return CRIMSON_ESTATE on the host without replaying; a new capture
    works again."""
    cfa, torch = env
    from crimsonflowsolver_amd.ensemble import Ensemble
    meshes, mu, R, states = _instances()
    ens = Ensemble(meshes, mu, R)
    y = np.concatenate([s[0] for s in states])
    ac = np.concatenate([s[1] for s in states])
    ty, tac, tx = _T(torch, y), _T(torch, ac), _T(torch, ens.x)
    r4 = torch.zeros((ens.N, 4), dtype=torch.float64, device="cuda")
    g = ens.capture(ty, tac, tx, r4, autotune=False)
    ens.set_parameters(mu=mu * 1.5)
    g.launch()
    got = r4.cpu().numpy().copy()
    ens.form.to_lhs(ty, tac, tx, r4, ens.lhs)
    assert np.array_equal(got, r4.cpu().numpy())
    # another instance count: the per-instance viscosity buffer is replaced
    ens.form.set_instances(np.zeros(ens.inst.size, np.uint16), np.array([0.05]))
    r4.fill_(7.0)
    with pytest.raises(cfa.CrimsonError) as ei:
        g.launch()
    assert ei.value.code == 8  # CRIMSON_ESTATE
    assert bool((r4 == 7.0).all())  # nothing was replayed
    ens.form.set_instances(None, np.zeros(0))
    with pytest.raises(cfa.CrimsonError):
        g.launch()
    g2 = ens.form.graph_to_lhs(ty, tac, tx, r4, ens.lhs)
    g2.launch()
    ens.form.to_lhs(ty, tac, tx, r4, ens.lhs)
    want = r4.cpu().numpy().copy()
    r4.zero_()
    g2.launch()
    assert np.array_equal(r4.cpu().numpy(), want)


def test_union_solve_matches_instances(env):
    cfa, torch = env
    from crimsonflowsolver_amd.ensemble import Ensemble
    meshes, mu, R, states = _instances()
    ens = Ensemble(meshes, mu, R)
    y = np.concatenate([s[0] for s in states])
    ac = np.concatenate([s[1] for s in states])
    r4 = torch.zeros((ens.N, 4), dtype=torch.float64, device="cuda")
    ens.form.to_lhs(_T(torch, y), _T(torch, ac), _T(torch, ens.x), r4, ens.lhs)
    ens.form.bc3res(r4, res4_layout=True)
    kw = dict(relTol=1e-6, dimKry=60, maxItr=10, relTolIn=[1e-4, 1e-4], maxItrIn=[10, 500])
    ls = cfa.memLS_LS_CREATE(cfa.LS_TYPE_NS, **kw)
    cfa.memLS_SOLVE_dev(ens.lhs, ls, 4, r4, res=ens.face_res())
    assert ls.RI["suc"] == 1
    xu = r4.cpu().numpy()
    for k, m in enumerate(meshes):
        colm, rowp = cfa.genadj(m.ien, m.numnp)
        from crimsonflowsolver_amd import mesh as M
        ibc, _ = M.dirichlet_codes(m)
        f = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, iBC=ibc, params=cfa.form_params(mu=float(mu[k])))
        lhs = cfa.memLS_LHS_CREATE(m.numnp, m.numnp, rowp.size, None, colm, rowp, 2)
        fixed = np.nonzero(ibc)[0]
        out = np.nonzero(m.outlet & (ibc == 0))[0]
        cfa.memLS_BC_CREATE(lhs, 1, fixed.size, 3, cfa.BC_TYPE_Dir, fixed + 1, np.zeros((3, fixed.size), order="F"))
        cfa.memLS_BC_CREATE(lhs, 2, out.size, 3, cfa.BC_TYPE_Neu, out + 1, np.asfortranarray(m.nabi()[out].T))
        rk = torch.zeros((m.numnp, 4), dtype=torch.float64, device="cuda")
        f.to_lhs(_T(torch, states[k][0]), _T(torch, states[k][1]), _T(torch, m.x), rk, lhs)
        f.bc3res(rk, res4_layout=True)
        lk = cfa.memLS_LS_CREATE(cfa.LS_TYPE_NS, **kw)
        p = ens.params
        cfa.memLS_SOLVE_dev(lhs, lk, 4, rk, res=np.array([0.0, R[k] * p.alfi * p.gami * p.Delt]))
        assert lk.RI["suc"] == 1
        xk = rk.cpu().numpy()
        err = np.abs(xu[ens.rows(k)] - xk).max() / np.abs(xk).max()
        print("instance", k, "union vs own solve", err)
        assert err <= 1e-4


def test_union_solve_many_coupled_faces(env):
    """18 instances: 18 coupled resistance faces in one memLS matrix (more
    than any fixed per-solve slot count), each face's ADD/PRE term must use
    its own face dot."""
    cfa, torch = env
    from crimsonflowsolver_amd.ensemble import Ensemble
    from crimsonflowsolver_amd import mesh as M
    K = 18
    meshes = [M.cylinder_mesh(4, 4, 4 + k % 5) for k in range(K)]
    mu = 0.03 + 0.002 * np.arange(K)
    R = np.linspace(800.0, 2000.0, K)
    states = [M.poiseuille_state(m, seed=90 + k) for k, m in enumerate(meshes)]
    ens = Ensemble(meshes, mu, R)
    y = np.concatenate([s[0] for s in states])
    ac = np.concatenate([s[1] for s in states])
    r4 = torch.zeros((ens.N, 4), dtype=torch.float64, device="cuda")
    ens.form.to_lhs(_T(torch, y), _T(torch, ac), _T(torch, ens.x), r4, ens.lhs)
    ens.form.bc3res(r4, res4_layout=True)
    kw = dict(relTol=1e-6, dimKry=60, maxItr=10, relTolIn=[1e-4, 1e-4], maxItrIn=[10, 500])
    ls = cfa.memLS_LS_CREATE(cfa.LS_TYPE_NS, **kw)
    cfa.memLS_SOLVE_dev(ens.lhs, ls, 4, r4, res=ens.face_res())
    assert ls.RI["suc"] == 1
    xu = r4.cpu().numpy()
    p = ens.params
    worst = 0.0
    for k, m in enumerate(meshes):
        colm, rowp = cfa.genadj(m.ien, m.numnp)
        ibc, _ = M.dirichlet_codes(m)
        f = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, iBC=ibc, params=cfa.form_params(mu=float(mu[k])))
        lhs = cfa.memLS_LHS_CREATE(m.numnp, m.numnp, rowp.size, None, colm, rowp, 2)
        fixed = np.nonzero(ibc)[0]
        out = np.nonzero(m.outlet & (ibc == 0))[0]
        cfa.memLS_BC_CREATE(lhs, 1, fixed.size, 3, cfa.BC_TYPE_Dir, fixed + 1, np.zeros((3, fixed.size), order="F"))
        cfa.memLS_BC_CREATE(lhs, 2, out.size, 3, cfa.BC_TYPE_Neu, out + 1, np.asfortranarray(m.nabi()[out].T))
        rk = torch.zeros((m.numnp, 4), dtype=torch.float64, device="cuda")
        f.to_lhs(_T(torch, states[k][0]), _T(torch, states[k][1]), _T(torch, m.x), rk, lhs)
        f.bc3res(rk, res4_layout=True)
        lk = cfa.memLS_LS_CREATE(cfa.LS_TYPE_NS, **kw)
        cfa.memLS_SOLVE_dev(lhs, lk, 4, rk, res=np.array([0.0, R[k] * p.alfi * p.gami * p.Delt]))
        xk = rk.cpu().numpy()
        err = np.abs(xu[ens.rows(k)] - xk).max() / np.abs(xk).max()
        worst = max(worst, err)
    print("18 coupled faces: worst union vs own solve", worst)
    assert worst <= 1e-4
