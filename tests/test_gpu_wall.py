"""GPU parity of the deformable wall (crimson_form_set_wall: e3bvar's wall
terms on the iBCB-bit-4 faces, e3b.f90:374-575, e3bvar.f90:462-1017)
against the oracle's restatement (cfo_boundary_wall, pinned in
test_wall_pins.py):

* the formation with the wall (membrane with a reference displacement,
  mass, tissue support and damping, ring tethering at both end rings):
  residual and lhsK (the wall tangent through bc3lhs and fillsparseI) and
  lhsP, FP64, normwise relative 1e-12; into the memLS planes (FP64 1e-12,
  FP32 tangent 1e-5);
* timesteps of the deformable cylinder (generalized alpha with the wall
  displacement's predictor, alpha level and corrector) against the oracle's
  timesteps, each iterated to a residual 1e-10 of its first: y, ac and the
  displacement within 1e-8 relative after 3 steps;
* a timestep captured with the wall (crimson_step_graph_create with the
  wall displacement's predictor, alpha level and corrector) equals the same
  timestep run entry point by entry point, bit for bit, over 3 replays;
* the entry points that do not carry the wall refuse it.
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


def _case(seed, grid=(6, 6, 10)):
    from crimsonflowsolver_amd import mesh as M
    m = M.cylinder_mesh(*grid)
    rng = np.random.default_rng(seed)
    y, ac = M.poiseuille_state(m, seed=seed)
    ac[:, :3] = rng.uniform(-50, 50, size=(m.numnp, 3))
    u = rng.normal(size=(m.numnp, 3)) * 1e-3
    uref = rng.normal(size=(m.numnp, 3)) * 1e-3
    tag = np.zeros(m.numnp, np.int32)
    tag[m.wall & (m.inlet | m.outlet)] = 1  # the wall's end rings
    return m, y, ac, u, uref, tag


def _oracle(O, m, y, ac, u, uref, tag, colm, rowp, bnd):
    p = O.make_params()
    rres, rK, rP = O.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, p)
    N = m.numnp
    rres = O.boundary_wall(*bnd, y, ac, u, m.x, rres, O.make_wall(p), uref=uref, tag=tag,
                           ibc=np.zeros(N, np.int32), bc=np.zeros((N, 6)), colm=colm, rowp=rowp, lhsK=rK)
    return rres, rK, rP


def test_wall_formation_matches_oracle(cfa, oracle_lib):
    import torch
    from crimsonflowsolver_amd import mesh as M
    m, y, ac, u, uref, tag = _case(41)
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    bnd = M.cylinder_boundary(m)
    rigid = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp)
    rigid.set_boundary(*bnd)
    res0, K0, P0 = rigid(y, ac, m.x)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, params=cfa.form_params(ideformwall=1))
    form.set_boundary(*bnd)
    with pytest.raises(cfa.CrimsonError):  # no wall yet
        form(y, ac, m.x)
    form.set_wall(cfa.wall_params(), nodetag=tag, uref=uref)
    with pytest.raises(cfa.CrimsonError):  # no displacement yet
        form(y, ac, m.x)
    du = torch.from_numpy(np.ascontiguousarray(u.T)).cuda()
    form.set_displacement(du)
    res, K, P = form(y, ac, m.x)
    rres, rK, rP = _oracle(oracle_lib, m, y, ac, u, uref, tag, colm, rowp, bnd)
    errs = {"res": _rel(res, rres), "lhsK": _rel(K, rK), "lhsP": _rel(P, rP)}
    print("deformable wall", errs, "wall share of res", _rel(res, res0), "of lhsK", _rel(K, K0))
    assert all(e <= TOL for e in errs.values()), errs
    assert _rel(res, res0) > 1e-3 and _rel(K, K0) > 1e-3  # the wall changed both
    assert np.array_equal(P, P0)
    # removing the boundary removes the wall; a new boundary needs a new wall
    form.set_boundary(*bnd)
    with pytest.raises(cfa.CrimsonError):
        form(y, ac, m.x)


@pytest.mark.parametrize("precision", [64, 32])
def test_wall_into_memls_planes(cfa, oracle_lib, precision):
    import torch
    from crimsonflowsolver_amd import mesh as M
    O = oracle_lib
    m, y, ac, u, uref, tag = _case(43, grid=(5, 5, 8))
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    bnd = M.cylinder_boundary(m)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp,
                              params=cfa.form_params(ideformwall=1, precision=precision))
    form.set_boundary(*bnd)
    form.set_wall(cfa.wall_params(), nodetag=tag, uref=uref)
    du = torch.from_numpy(np.ascontiguousarray(u.T)).cuda()
    form.set_displacement(du)
    lhs = cfa.memLS_LHS_CREATE(m.numnp, m.numnp, rowp.size, None, colm, rowp, 1)  # FP32 planes from the form
    T = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, order="F").T)).cuda()
    res4 = torch.zeros((m.numnp, 4), dtype=torch.float64, device="cuda")
    form.to_lhs(T(y), T(ac), T(m.x), res4, lhs)
    form.sync()
    rres, rK, rP = _oracle(O, m, y, ac, u, uref, tag, colm, rowp, bnd)
    val = lhs.get_val(4)
    want = O.pack_val4(colm, rowp, rK, rP)
    e_res = _rel(res4.cpu().numpy(), rres)
    e_val = _rel(val, want)
    print(f"FP{precision}: Res4 {e_res:.2e}, Val4 {e_val:.2e}")
    assert e_res <= TOL
    assert e_val <= (TOL if precision == 64 else 1e-5)


def test_wall_on_the_reference_basic_case(cfa, oracle_lib):
    """The reference's own deformable-wall cylinder (testbin/mainTests/basic:
    its mesh, boundary codes, node tags with the 28 ring nodes, restart state
    and reference displacement, its solver.inp's material, time integration
    and wall values), with a seeded wall displacement."""
    import os
    import torch
    O = oracle_lib
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "basic_case.npz"))
    x, ien, ienb, iBCB, BCB = z["x"], z["ien"], z["ienb"], z["iBCB"], z["BCB"]
    y, ac, uref, tag = z["y"], z["ac"], z["uref"], z["nodetag"]
    N = x.shape[0]
    rng = np.random.default_rng(8)
    u = rng.normal(size=(N, 3)) * 1e-2
    colm, rowp = cfa.genadj(ien, N)
    form = cfa.ElmGMRInterior(ien, colm, rowp, N,
                              params=cfa.form_params(rho=0.00106, mu=0.004, rhoinf=0.0, ideformwall=1))
    form.set_boundary(ienb, iBCB, BCB)
    form.set_wall(cfa.wall_params(rhovw=0.00106), nodetag=tag, uref=uref)
    form.set_displacement(torch.from_numpy(np.ascontiguousarray(u.T)).cuda())
    res, K, P = form(y, ac, x)
    p = O.make_params(rho=0.00106, mu=0.004, rhoinf=0.0)
    rres, rK, rP = O.elmgmr_interior(ien, y, ac, x, colm, rowp, p)
    rres = O.boundary_wall(ienb, iBCB, BCB, y, ac, u, x, rres, O.make_wall(p, rhovw=0.00106), uref=uref, tag=tag,
                           rho=0.00106, mu=0.004, ibc=np.zeros(N, np.int32), bc=np.zeros((N, 6)), colm=colm,
                           rowp=rowp, lhsK=rK)
    errs = {"res": _rel(res, rres), "lhsK": _rel(K, rK), "lhsP": _rel(P, rP)}
    print("reference basic case, deformable wall", errs, "ring nodes", int(tag.sum()))
    assert all(e <= TOL for e in errs.values()), errs


def test_wall_timesteps_match_oracle(cfa, oracle_lib):
    import os
    import sys
    import torch
    sys.path.insert(0, os.path.dirname(__file__))
    from oracle_steps import OracleCylinder
    from crimsonflowsolver_amd import mesh as M
    from crimsonflowsolver_amd.timestep import ResistanceCylinder, ns_solver
    O = oracle_lib
    m = M.cylinder_mesh(5, 5, 8)
    rng = np.random.default_rng(5)
    uref = rng.normal(size=(m.numnp, 3)) * 1e-3
    tag = np.zeros(m.numnp, np.int32)
    tag[m.wall & (m.inlet | m.outlet)] = 1
    ls = ns_solver(relTol=1e-6, maxItr=10, relTolIn=(1e-3, 1e-3), maxItrIn=(10, 500))
    c = ResistanceCylinder(m, wall=cfa.wall_params(), nodetag=tag, uref=uref, ls=ls)
    oc = OracleCylinder(O, m, wall=O.make_wall(O.make_params()), nodetag=tag, uref=uref)
    y, ac = c.state()
    u = torch.zeros((3, m.numnp), dtype=torch.float64, device="cuda")
    Y, A = (t.cpu().numpy().T.copy(order="F") for t in (y, ac))
    U = np.zeros((m.numnp, 3))
    p = c.params
    for n in range(3):
        t = (n + 1) * p.Delt
        y, ac, u, gn = c.step_wall(y, ac, u, t, tol=1e-10, unit_rhs=True, max_iters=60)
        Y, A, U, on = oc.step(Y, A, t, tol=1e-10, max_iters=60, uold=U)
        print(f"step {n + 1}: GPU {len(gn) - 1} iterations to {gn[-1] / gn[0]:.1e}, oracle {len(on) - 1} to "
              f"{on[-1] / on[0]:.1e}", flush=True)
        assert gn[-1] <= 1e-10 * gn[0] and on[-1] <= 1e-10 * on[0]
    gy, ga, gu = (t.cpu().numpy().T for t in (y, ac, u))
    wn = oc.wall_nodes
    errs = {"y": _rel(gy, Y), "ac": _rel(ga, A), "u": _rel(gu[wn], U[wn])}
    print("after 3 steps", errs, "wall displacement max", np.abs(U[wn]).max())
    assert np.abs(U[wn]).max() > 0
    assert all(e <= 1e-8 for e in errs.values()), errs


def test_wall_graph_timestep_equals_eager(cfa):
    import torch
    from crimsonflowsolver_amd import mesh as M
    from crimsonflowsolver_amd.timestep import ResistanceCylinder, ns_solver
    m = M.cylinder_mesh(5, 5, 8)
    rng = np.random.default_rng(12)
    uref = rng.normal(size=(m.numnp, 3)) * 1e-3
    tag = np.zeros(m.numnp, np.int32)
    tag[m.wall & (m.inlet | m.outlet)] = 1
    mk = lambda: ns_solver(relTol=1e-6, dimKry=40, maxItr=6, relTolIn=(0.01, 0.2), maxItrIn=(1, 200))
    c = ResistanceCylinder(m, wall=cfa.wall_params(), nodetag=tag, uref=uref, ls=mk())
    y0, ac0 = c.state()
    u0 = torch.zeros((3, m.numnp), dtype=torch.float64, device="cuda")
    p = c.params
    states = []
    y, ac, u = y0.clone(), ac0.clone(), u0.clone()
    for n in range(3):
        y, ac, u, _ = c.step_wall(y, ac, u, (n + 1) * p.Delt, iters=3)
        states.append((y.clone(), ac.clone(), u.clone()))
    ls_g = mk()
    gy0, gac0, gu0 = y0.clone(), ac0.clone(), u0.clone()
    gy, gac, gya, gaca = (torch.empty_like(y0) for _ in range(4))
    gu = u0.clone()
    res4 = torch.zeros((c.N, 4), dtype=torch.float64, device="cuda")
    bc = c.bc(p.Delt)
    c.form.set_wall_state(gu0, gu)
    g = c.form.step_graph(c.lhs, ls_g, gy0, gac0, gy, gac, gya, gaca, c.dx, res4, c.dibc, bc,
                          surfaces=[(c.dglob, c.dnabi, c.R, 0.0)], faceres=c.faceres, nonlinear=3)
    for n in range(3):
        bc.copy_(c.bc((n + 1) * p.Delt))
        torch.cuda.synchronize()
        g.launch()
        ok = [torch.equal(gy, states[n][0]), torch.equal(gac, states[n][1]), torch.equal(gu, states[n][2])]
        print(f"timestep {n + 1}: y, ac, u bitwise equal {ok}; displacement max {float(gu.abs().max()):.3e}",
              flush=True)
        assert all(ok), n
        assert torch.equal(gu0, gu)  # uold = u at the end of the captured step


def test_entry_points_without_the_wall_refuse_it(cfa):
    import torch
    from crimsonflowsolver_amd import mesh as M
    from crimsonflowsolver_amd.timestep import ResistanceCylinder, ns_solver
    m = M.cylinder_mesh(4, 4, 5)
    c = ResistanceCylinder(m, wall=cfa.wall_params(iringsupp=0, iringdamp=0), ls=ns_solver())
    y, ac = c.state()
    w = torch.zeros_like(y)
    with pytest.raises(cfa.CrimsonError) as e:
        c.form.vjp(y, ac, c.dx, w, torch.empty_like(y))
    assert e.value.code == 2  # CRIMSON_EUNSUPPORTED
    with pytest.raises(ValueError):
        c.step(y, ac, c.params.Delt)
