"""Adjoint of the implicit generalized-alpha timestep (adjoint.StepAdjoint) on
the GPU, checked against the oracle.

* crimson_lhs_transpose_dev: every block of the transposed matrix equals the
  transpose of its mirror block (exact).
* crimson_form_vjp_ac_dev: the acceleration's cotangent against central
  differences of the oracle residual in ac (h 1e-6 relative; 1e-6).
* crimson_elmpvsq_vjp_dev against the closed form of the outlet term.
* the step operator F_z^T (mask, generalized-alpha scaling, ElmpvsQ term)
  against central differences of the oracle's step residual F in z.
* the gradient of J = <cy, y_N> + <ca, ac_N> over 6 steps (checkpoints every
  2) and over 50 steps (checkpoints every 10) on a small cylinder, with
  respect to mu, the outlet resistance R and the initial state, against
  central differences of J computed by oracle-driven steps (each step
  iterated until its residual is 1e-12 of its first, or below 1e-10;
  relative 1e-5).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# steps are iterated until the residual norm is 1e-12 of its first value or
# below this floor (the residual's rounding level on these meshes, where the
# later steps start from residuals of order 10)
ATOL = 1e-10


@pytest.fixture(scope="module")
def env():
    import torch
    import crimsonflowsolver_amd as cfa
    if cfa.device_count() < 1:
        pytest.fail("no GPU visible: the HIP path must run here")
    return cfa, torch


def _case(env, grid=(4, 4, 8), mu=0.04, R=1333.0):
    cfa, torch = env
    from crimsonflowsolver_amd import mesh as M
    from crimsonflowsolver_amd.timestep import ResistanceCylinder, ns_solver
    m = M.cylinder_mesh(*grid)
    ls = ns_solver(relTol=1e-6, maxItr=10, relTolIn=(1e-3, 1e-3), maxItrIn=(10, 500))
    return ResistanceCylinder(m, R=R, params=cfa.form_params(mu=mu), ls=ls, transpose=True)


def _H(t):
    return t.cpu().numpy()


def test_transposed_matrix_blocks(env):
    cfa, torch = env
    c = _case(env)
    y, ac = c.state()
    c.residual(y, ac, y, ac)
    c.lhsT.transpose_from(c.lhs)
    V, Vt = c.lhs.get_val(4), c.lhsT.get_val(4)
    rowp, colm = c.rowp - 1, c.colm - 1
    rows = np.repeat(np.arange(c.N), np.diff(colm))
    key = {(int(i), int(j)): k for k, (i, j) in enumerate(zip(rows, rowp))}
    tr = np.array([key[(int(j), int(i))] for i, j in zip(rows, rowp)])
    want = V[:, tr].reshape(4, 4, -1).transpose(1, 0, 2).reshape(16, -1)
    assert np.array_equal(Vt, want)


def test_vjp_acceleration_against_oracle_fd(env, oracle_lib):
    cfa, torch = env
    O = oracle_lib
    from crimsonflowsolver_amd import mesh as M
    m = M.cylinder_mesh(5, 5, 7)
    rng = np.random.default_rng(81)
    y, ac = M.poiseuille_state(m, seed=81)
    ac[:, :3] = rng.uniform(-40, 40, size=(m.numnp, 3))
    colm, rowp = O.genadj(m.ien, m.numnp)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, order="F").T)).cuda()
    w = rng.normal(size=(m.numnp, 4))
    ty, tac, tx, tw = T(y), T(ac), T(m.x), T(w)
    yb, ab, yb0 = torch.empty_like(ty), torch.empty_like(ty), torch.empty_like(ty)
    mub = form.vjp_ac(ty, tac, tx, tw, yb, ab)
    mub0 = form.vjp(ty, tac, tx, tw, yb0)
    assert torch.equal(yb, yb0) and mub == mub0  # the state part is unchanged
    ab = _H(ab).T
    assert np.all(ab[:, 3] == 0.0)

    def J(aa):
        r, _, _ = O.elmgmr_interior(m.ien, y, aa, m.x, colm, rowp, O.make_params(lhs=0))
        return float(np.sum(w * r))

    for k in range(3):
        d = np.zeros_like(ac)
        d[:, :3] = rng.normal(size=(m.numnp, 3))
        h = 1e-6 * np.abs(ac).max() / np.abs(d).max()
        fd = (J(ac + h * d) - J(ac - h * d)) / (2 * h)
        an = float(np.sum(ab * d))
        print("direction", k, "fd", fd, "vjp", an)
        assert abs(fd - an) <= 1e-6 * abs(fd)


def test_elmpvsq_vjp_closed_form(env):
    cfa, torch = env
    c = _case(env)
    rng = np.random.default_rng(5)
    y = torch.from_numpy(rng.normal(size=(4, c.N))).cuda()
    w4 = torch.from_numpy(rng.normal(size=(c.N, 4))).cuda()
    ybar = torch.from_numpy(rng.normal(size=(4, c.N))).cuda()
    y0 = _H(ybar).copy()
    cbar = torch.zeros(2, dtype=torch.float64, device="cuda")
    R = 1500.0
    c.form.elmpvsq_vjp(c.dglob, c.dnabi, R, y, w4, ybar, cbar)
    c.form.sync()
    nb, g = c.nabi[c.out], c.out
    s = float(np.sum(_H(w4)[g, :3] * nb))
    Q = float(np.sum(_H(y)[:3, g].T * nb))
    want = y0.copy()
    want[:3, g] += (-R * s * nb).T
    got = _H(ybar)
    assert np.allclose(got, want, rtol=1e-14, atol=1e-12 * np.abs(want).max())
    cb = _H(cbar)
    assert abs(cb[0] + Q * s) <= 1e-13 * abs(Q * s) and abs(cb[1] + s) <= 1e-13 * abs(s)


def test_step_operator_against_oracle_fd(env, oracle_lib):
    """lam^T (F(z + h dz) - F(z - h dz)) / 2h with the oracle's F against
    (F_z^T lam) . dz, at a partly converged iterate of one step."""
    cfa, torch = env
    import sys
    import os
    sys.path.insert(0, os.path.dirname(__file__))
    from oracle_steps import OracleCylinder
    from crimsonflowsolver_amd.adjoint import StepAdjoint
    c = _case(env)
    oc = OracleCylinder(oracle_lib, c.mesh)
    p = c.params
    yn, acn = c.state()
    y1, ac1, _ = c.step(yn, acn, p.Delt, iters=2)
    adj = StepAdjoint(c)
    ya, aca = torch.empty_like(yn), torch.empty_like(acn)
    c.form.yalpha(yn, acn, y1, ac1, ya, aca)
    c.form.sync()
    rng = np.random.default_rng(9)
    mask = _H(c.mask4)
    lam = rng.normal(size=(c.N, 4)) * mask
    dz = rng.normal(size=(c.N, 4)) * mask * np.array([1.0, 1.0, 1.0, 30.0])
    out = _H(adj._op(ya, aca, torch.from_numpy(lam).cuda())[0])
    an = float(np.sum(out * dz))
    Yn, An, Y1, A1 = (_H(t).T.copy(order="F") for t in (yn, acn, y1, ac1))
    gdt = p.gami * p.Delt

    def F(h):
        Y, A = Y1.copy(order="F"), A1.copy(order="F")
        A[:, :3] += h * dz[:, :3]
        Y[:, :3] += gdt * h * dz[:, :3]
        Y[:, 3] += p.alfi * gdt * h * dz[:, 3]
        r, _, _ = oc.residual(Yn, An, Y, A)
        return float(np.sum(lam * r))

    h = 1e-6 * np.abs(A1).max() / np.abs(dz).max()
    fd = (F(h) - F(-h)) / (2 * h)
    print("fd", fd, "adjoint operator", an)
    assert abs(fd - an) <= 1e-6 * abs(fd)


def _oracle_J(O, mesh, y0, ac0, nsteps, cy, ca, mu=0.04, R=1333.0):
    import sys
    import os
    sys.path.insert(0, os.path.dirname(__file__))
    from oracle_steps import OracleCylinder
    oc = OracleCylinder(O, mesh, R=R, mu=mu)
    y, ac = y0, ac0
    for n in range(nsteps):
        y, ac, norms = oc.step(y, ac, (n + 1) * oc.p.Delt, tol=1e-12, atol=ATOL)
        assert norms[-1] <= max(1e-12 * norms[0], ATOL)
    return float(np.sum(cy * y) + np.sum(ca * ac))


def _adjoint_against_fd(env, O, nsteps, every, checks):
    cfa, torch = env
    from crimsonflowsolver_amd.adjoint import StepAdjoint
    c = _case(env)
    rng = np.random.default_rng(nsteps)
    y0, ac0 = c.state()
    Y0, A0 = _H(y0).T.copy(order="F"), _H(ac0).T.copy(order="F")
    cy, ca = rng.normal(size=(c.N, 4)), rng.normal(size=(c.N, 4))
    ca[:, 3] = 0.0
    adj = StepAdjoint(c, tol=1e-11)
    kw = dict(tol=1e-12, unit_rhs=True, atol=ATOL)
    J, yb, ab, mub, rb = adj.gradient(y0, ac0, 0.0, nsteps, every, c.T(cy), c.T(ca), step_kw=kw)
    its = [s[0] for s in adj.stats]
    print(f"{nsteps} steps: J {J}, transpose-solve iterations {min(its)}-{max(its)}, "
          f"worst residual {max(s[1] for s in adj.stats):.1e}", flush=True)
    Jo = _oracle_J(O, c.mesh, Y0, A0, nsteps, cy, ca)
    print("J gpu", J, "oracle", Jo, flush=True)
    assert abs(J - Jo) <= 1e-10 * abs(Jo)
    yb = _H(yb).T
    if "mu" in checks:
        h = 1e-5 * 0.04
        fd = (_oracle_J(O, c.mesh, Y0, A0, nsteps, cy, ca, mu=0.04 + h) -
              _oracle_J(O, c.mesh, Y0, A0, nsteps, cy, ca, mu=0.04 - h)) / (2 * h)
        print("mu: fd", fd, "adjoint", mub, flush=True)
        assert abs(fd - mub) <= 1e-5 * abs(fd)
    if "R" in checks:
        h = 1e-5 * 1333.0
        fd = (_oracle_J(O, c.mesh, Y0, A0, nsteps, cy, ca, R=1333.0 + h) -
              _oracle_J(O, c.mesh, Y0, A0, nsteps, cy, ca, R=1333.0 - h)) / (2 * h)
        print("R: fd", fd, "adjoint", rb, flush=True)
        assert abs(fd - rb) <= 1e-5 * abs(fd)
    if "y0" in checks:
        d = rng.normal(size=(c.N, 4)) * np.array([1.0, 1.0, 1.0, 30.0])
        h = 1e-5 * np.abs(Y0).max() / np.abs(d).max()
        fd = (_oracle_J(O, c.mesh, Y0 + h * d, A0, nsteps, cy, ca) -
              _oracle_J(O, c.mesh, Y0 - h * d, A0, nsteps, cy, ca)) / (2 * h)
        an = float(np.sum(yb * d))
        print("y0: fd", fd, "adjoint", an, flush=True)
        assert abs(fd - an) <= 1e-5 * abs(fd)


def test_step_adjoint_6_steps_against_oracle_fd(env, oracle_lib):
    _adjoint_against_fd(env, oracle_lib, 6, 2, ("mu", "R", "y0"))


def test_step_adjoint_50_steps_checkpointed_against_oracle_fd(env, oracle_lib):
    _adjoint_against_fd(env, oracle_lib, 50, 10, ("mu",))
