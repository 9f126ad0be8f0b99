"""Device-driven NS solves and whole timesteps captured in HIP graphs
(crimson_ls_graph_create, crimson_step_graph_create).

* A captured NS solve against crimson_ls_solve_dev on the same matrix and
  right-hand side, FP64 and FP32 matrices, with a coupled resistance face:
  the same solution bit for bit and the same statistics; three replays, each
  after eager work on the device, each checked.
* A captured timestep (3 nonlinear iterations) of the resistance cylinder
  against the same timestep run entry point by entry point: bitwise equal
  states after each of three replays (the inflow rewritten between them).
* The ensemble timestep: three instances on uneven cylinders as one union
  mesh, one launch per timestep, against each instance's own timestep driven
  by the oracle (the same steps and solver settings, NS tolerance 1e-6):
  states within the solver's tolerance (the union's Krylov space spans all
  instances, so the iterates differ; relative 1e-5 of the state after 3
  timesteps).
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


def _case(env, grid=(6, 6, 10), precision=64):
    cfa, torch = env
    from crimsonflowsolver_amd import mesh as M
    from crimsonflowsolver_amd.timestep import ResistanceCylinder, ns_solver
    m = M.cylinder_mesh(*grid)
    ls = ns_solver(relTol=1e-6, dimKry=40, maxItr=6, relTolIn=(0.01, 0.2), maxItrIn=(1, 200))
    return ResistanceCylinder(m, params=cfa.form_params(precision=precision), ls=ls)


@pytest.mark.parametrize("precision", [64, 32])
def test_graph_solve_equals_eager_solve(env, precision):
    cfa, torch = env
    from crimsonflowsolver_amd.timestep import ns_solver
    c = _case(env, precision=precision)
    y, ac = c.state()
    c.residual(y, ac, y, ac)
    rhs = c.res4.clone()
    eager = rhs.clone()
    cfa.memLS_SOLVE_dev(c.lhs, c.ls, 4, eager, res=c.faceres)
    want = {s: c.ls.stats(s) for s in ("RI", "GM", "CG")}
    print(f"FP{precision} eager: RI {want['RI']['itr']:.0f} GM {want['GM']['itr']:.0f} CG {want['CG']['itr']:.0f} "
          f"suc {want['RI']['suc']:.0f}", flush=True)
    ls_g = ns_solver(relTol=1e-6, dimKry=40, maxItr=6, relTolIn=(0.01, 0.2), maxItrIn=(1, 200))
    buf = torch.empty_like(rhs)
    g = cfa.memLS_SOLVE_graph(c.lhs, ls_g, 4, buf, res=c.faceres)
    for rep in range(3):
        buf.copy_(rhs)
        torch.cuda.synchronize()
        g.launch()
        assert torch.equal(buf, eager), rep
        for s in ("RI", "GM", "CG"):
            got = ls_g.stats(s)
            for k in ("iNorm", "fNorm", "itr", "suc"):
                assert got[k] == want[s][k], (rep, s, k, got[k], want[s][k])
        # eager work on the device between replays
        tmp = (buf * 2.0).sum()
        torch.cuda.synchronize()
        del tmp


def test_graph_timestep_equals_eager_timestep(env):
    cfa, torch = env
    from crimsonflowsolver_amd.timestep import ns_solver
    c = _case(env)
    y0, ac0 = c.state()
    p = c.params
    # eager: 3 timesteps of 3 nonlinear iterations
    states = []
    yold, acold = y0.clone(), ac0.clone()
    for n in range(3):
        y, ac, _ = c.step(yold, acold, (n + 1) * p.Delt, iters=3)
        states.append((y.clone(), ac.clone()))
        yold, acold = y, ac
    # captured
    ls_g = ns_solver(relTol=1e-6, dimKry=40, maxItr=6, relTolIn=(0.01, 0.2), maxItrIn=(1, 200))
    gy0, gac0 = y0.clone(), ac0.clone()
    gy, gac, gya, gaca = (torch.empty_like(y0) for _ in range(4))
    res4 = torch.zeros((c.N, 4), dtype=torch.float64, device="cuda")
    bc = c.bc(p.Delt)
    g = c.form.step_graph(c.lhs, ls_g, gy0, gac0, gy, gac, gya, gaca, c.dx, res4, c.dibc, bc,
                          surfaces=[(c.dglob, c.dnabi, c.R, 0.0)], faceres=c.faceres, nonlinear=3)
    for n in range(3):
        bc.copy_(c.bc((n + 1) * p.Delt))
        torch.cuda.synchronize()
        g.launch()
        print(f"timestep {n + 1}: RI {ls_g.RI['itr']:.0f} CG {ls_g.CG['itr']:.0f}", flush=True)
        assert torch.equal(gy, states[n][0]) and torch.equal(gac, states[n][1]), n
        assert torch.equal(gy0, gy) and torch.equal(gac0, gac)


def test_ensemble_graph_timestep_against_oracle_instances(env, oracle_lib):
    cfa, torch = env
    import os
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    from oracle_steps import OracleCylinder
    from crimsonflowsolver_amd import mesh as M
    from crimsonflowsolver_amd.ensemble import Ensemble
    from crimsonflowsolver_amd.timestep import ns_solver
    meshes = [M.cylinder_mesh(5, 5, 6), M.cylinder_mesh(6, 6, 9, length=7.0), M.cylinder_mesh(4, 4, 12)]
    mu = np.array([0.03, 0.04, 0.06])
    R = np.array([800.0, 1333.0, 2000.0])
    ens = Ensemble(meshes, mu, R, boundary=True)
    p = ens.params
    T = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, order="F").T)).cuda()
    st = [M.poiseuille_state(mm, seed=70 + k) for k, mm in enumerate(meshes)]
    yold = T(np.concatenate([s[0] for s in st]))
    acold = T(np.concatenate([s[1] for s in st]))
    y, ac, ya, aca = (torch.empty_like(yold) for _ in range(4))
    x = T(ens.x)
    res4 = torch.zeros((ens.N, 4), dtype=torch.float64, device="cuda")
    ibc = torch.from_numpy(ens.iBC).cuda()
    inflow = 10.0
    bc_h = np.zeros((ens.N, 7), order="F")
    for k, mm in enumerate(meshes):
        rr = mm.x[:, 0] ** 2 + mm.x[:, 1] ** 2
        inl = mm.inlet & ~mm.wall
        col = np.zeros(mm.numnp)
        col[inl] = inflow * 2.0 * np.clip(1.0 - rr[inl] / mm.radius ** 2, 0.0, None)
        bc_h[ens.rows(k), 4] = col
    bc = T(bc_h)
    surfaces = [(torch.from_numpy(ens.outlets[k].astype(np.int32)).cuda(),
                 torch.from_numpy(np.ascontiguousarray(ens.nabi[k])).cuda(), R[k], 0.0) for k in range(ens.K)]
    # the oracle stepper's solver settings (OracleCylinder)
    tol = 1e-6
    ls = ns_solver(relTol=tol, dimKry=100, maxItr=10, relTolIn=(1e-3, 1e-3), maxItrIn=(10, 500))
    g = ens.form.step_graph(ens.lhs, ls, yold, acold, y, ac, ya, aca, x, res4, ibc, bc, surfaces=surfaces,
                            faceres=ens.face_res(), nonlinear=3)
    oc = [OracleCylinder(oracle_lib, mm, R=R[k], mu=mu[k], relTol=tol) for k, mm in enumerate(meshes)]
    ost = [(s[0].copy(order="F"), s[1].copy(order="F")) for s in st]
    for n in range(3):
        g.launch()
        print(f"timestep {n + 1}: union RI {ls.RI['itr']:.0f} GM {ls.GM['itr']:.0f} CG {ls.CG['itr']:.0f}",
              flush=True)
        for k in range(ens.K):
            yk, ak, _ = oc[k].step(*ost[k], (n + 1) * p.Delt, tol=0.0, max_iters=3)
            ost[k] = (yk, ak)
    got = y.cpu().numpy().T
    for k in range(ens.K):
        gk, ok = got[ens.rows(k)], ost[k][0]
        err_u = np.abs(gk[:, :3] - ok[:, :3]).max() / np.abs(ok[:, :3]).max()
        err_p = np.abs(gk[:, 3] - ok[:, 3]).max() / np.abs(ok[:, 3]).max()
        print(f"instance {k}: rel err u {err_u:.2e} p {err_p:.2e}", flush=True)
        assert err_u <= 1e-5 and err_p <= 1e-5, (k, err_u, err_p)


def test_graph_refuses_a_solve_past_its_launch_budget(env):
    """A device-driven solve whose GMRES call would need more launches than a
    captured segment may hold is refused before anything is captured."""
    cfa, torch = env
    from crimsonflowsolver_amd.timestep import ns_solver
    c = _case(env, grid=(4, 4, 6))
    y, ac = c.state()
    c.residual(y, ac, y, ac)
    big = ns_solver(relTol=1e-6, dimKry=250, maxItr=2, relTolIn=(0.01, 0.2), maxItrIn=(20, 100))
    buf = c.res4.clone()
    with pytest.raises(cfa.CrimsonError) as e:
        cfa.memLS_SOLVE_graph(c.lhs, big, 4, buf, res=c.faceres)
    assert e.value.code == 2 and "launches" in str(e.value)
