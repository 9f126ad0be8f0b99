"""Adjoint (VJP) of the interior residual on the GPU, through the C ABI
(crimson_form_vjp_dev) and the torch.autograd harness.

Checks:
  * central differences of the oracle's residual (the CPU restatement) in a
    random direction of y and in mu, h = 1e-6 relative: the VJP matches to
    1e-6 relative (the FD truncation error at this h is ~1e-10);
  * the assembled tangent transposed where the tangent is the exact
    Jacobian: the momentum rows' velocity block at u = 0 and the continuity
    rows' pressure block (the PSPG matrix C) at a generic state, to 1e-10
    relative;
  * the ElmpvsQ resistance term's gradient in u and R against FD;
  * the checkpointed multi-step adjoint equals the plain one and FD.
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


def _state(grid, seed, generic=True):
    from crimsonflowsolver_amd import mesh as M
    m = M.cylinder_mesh(*grid)
    y, ac = M.poiseuille_state(m, seed=seed, noise=0.2)
    rng = np.random.default_rng(seed)
    if generic:
        ac[:, :3] = rng.uniform(-50, 50, size=(m.numnp, 3))
    return m, y, ac, rng


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _vjp(env, form, m, y, ac, w):
    cfa, torch = env
    F = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, order="F").T)).cuda()
    ybar = torch.empty_like(F(y))
    mub = form.vjp(F(y), F(ac), F(m.x), F(w), ybar)
    return ybar.cpu().numpy().T, mub


def test_vjp_matches_central_differences(env, oracle_lib):
    cfa, torch = env
    m, y, ac, rng = _state((5, 5, 7), 41)
    colm, rowp = oracle_lib.genadj(m.ien, m.numnp)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp)
    w = np.asfortranarray(rng.normal(size=(m.numnp, 4)))
    ybar, mub = _vjp(env, form, m, y, ac, w)

    def J(yy, mu=0.04):
        r, _, _ = oracle_lib.elmgmr_interior(m.ien, yy, ac, m.x, colm, rowp, oracle_lib.make_params(mu=mu, lhs=0))
        return float(np.sum(w * r))

    for k in range(3):
        d = np.asfortranarray(rng.normal(size=y.shape) * np.array([1.0, 1.0, 1.0, 30.0]))
        h = 1e-6 * np.abs(y).max() / np.abs(d).max()
        fd = (J(y + h * d) - J(y - h * d)) / (2 * h)
        an = float(np.sum(ybar * d))
        print("direction", k, "fd", fd, "vjp", an)
        assert abs(fd - an) <= 1e-6 * max(abs(fd), 1e-300)
    hm = 1e-6 * 0.04
    fdm = (J(y, 0.04 + hm) - J(y, 0.04 - hm)) / (2 * hm)
    print("mu: fd", fdm, "vjp", mub)
    assert abs(fdm - mub) <= 1e-6 * abs(fdm)


def test_vjp_against_tangent_transpose(env, oracle_lib):
    cfa, torch = env
    from test_oracle_pins import _dense_blocks, _p1_mass
    p = oracle_lib.make_params()
    lf = p.alfi * p.gami * p.Delt
    # momentum rows, velocity columns at u = 0, constant p, a = 0: L = 0 and
    # every path through q or tau is multiplied by u or L, so the
    # tangent is the exact Jacobian there
    m, y, ac, rng = _state((4, 4, 5), 42, generic=False)
    y[:, :3] = 0.0
    y[:, 3] = 50.0
    ac[:] = 0.0
    colm, rowp = oracle_lib.genadj(m.ien, m.numnp)
    n = m.numnp
    _, lhsK, lhsP = oracle_lib.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, p)
    K, Gp, D, C = _dense_blocks(colm, rowp, lhsK, lhsP, n)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, n)
    w = np.zeros((n, 4), order="F")
    w[:, :3] = rng.normal(size=(n, 3))
    ybar, _ = _vjp(env, form, m, y, ac, w)
    # d res_mom / d u = (alpha_m rho M delta - K) / lf  (K^{ij} of equation i, variable j)
    Mm = p.almi * p.rho * _p1_mass(m)
    Kd = K.reshape(3 * n, 3 * n)  # [eq node, eq comp][var node, var comp] flattened as (n,3,n,3)
    dR = -Kd.copy()
    for c in range(3):
        dR[c::3, c::3] += Mm
    want = (dR / lf).T @ w[:, :3].reshape(-1)
    got = ybar[:, :3].reshape(-1)
    err = np.abs(got - want).max() / np.abs(want).max()
    print("velocity block err", err)
    assert err <= 1e-10
    # continuity rows, pressure columns at a generic state: d res_cont / d p = -C / lf
    m, y, ac, rng = _state((4, 4, 5), 43)
    colm, rowp = oracle_lib.genadj(m.ien, m.numnp)
    n = m.numnp
    _, lhsK, lhsP = oracle_lib.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, p)
    K, Gp, D, C = _dense_blocks(colm, rowp, lhsK, lhsP, n)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, n)
    w = np.zeros((n, 4), order="F")
    w[:, 3] = rng.normal(size=n)
    ybar, _ = _vjp(env, form, m, y, ac, w)
    want = -(C.T @ w[:, 3]) / lf
    err = np.abs(ybar[:, 3] - want).max() / np.abs(want).max()
    print("pressure block err", err)
    assert err <= 1e-10


def test_autograd_harness_and_resistance(env, oracle_lib):
    cfa, torch = env
    from crimsonflowsolver_amd import adjoint as A
    m, y, ac, rng = _state((5, 5, 6), 44)
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp)
    yt = _dev(torch, y).requires_grad_(True)
    mu = torch.tensor(0.04, dtype=torch.float64, device="cuda", requires_grad=True)
    R = torch.tensor(1333.0, dtype=torch.float64, device="cuda", requires_grad=True)
    out = np.nonzero(m.outlet)[0]
    glob = torch.from_numpy(out).cuda()
    nabi = _dev(torch, m.nabi()[out])
    w = _dev(torch, rng.normal(size=(m.numnp, 4)))
    res = A.interior_residual(form, yt, mu, _dev(torch, ac), _dev(torch, m.x)) + A.resistance_term(yt, R, glob, nabi)
    L = (w * res).sum()
    gy, gmu, gR = torch.autograd.grad(L, (yt, mu, R))
    # the interior part equals the direct VJP; the resistance part is -R (w.NABI) NABI and -Q (w.NABI)
    ybar, mub = _vjp(env, form, m, y, ac, w.cpu().numpy())
    Q = float((m.nabi()[out] * y[out, :3]).sum())
    wn = float((w.cpu().numpy()[out, :3] * m.nabi()[out]).sum())
    add = np.zeros_like(ybar)
    add[out, :3] = -1333.0 * wn * m.nabi()[out]
    np.testing.assert_allclose(gy.cpu().numpy(), ybar + add, rtol=0, atol=1e-12 * np.abs(ybar).max())
    assert abs(float(gmu) - mub) <= 1e-12 * abs(mub)
    assert abs(float(gR) - (-Q * wn)) <= 1e-12 * abs(Q * wn)


def test_checkpointed_multistep_adjoint(env, oracle_lib):
    """Four explicit pseudo-time steps y <- y - alpha R(y; mu): the
    checkpointed reverse sweep (a checkpoint every 2 steps) equals the plain
    one (every 4) and central differences of the whole recurrence."""
    cfa, torch = env
    from crimsonflowsolver_amd import adjoint as A
    m, y, ac, rng = _state((4, 4, 6), 45)
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    forms = {}

    def form_mu(mu):
        key = round(mu, 18)
        if key not in forms:
            forms[key] = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, params=cfa.form_params(mu=mu))
        return forms[key]

    act, xt = _dev(torch, ac), _dev(torch, m.x)
    r0 = A.interior_residual(form_mu(0.04), _dev(torch, y), torch.tensor(0.04, dtype=torch.float64, device="cuda"),
                             act, xt)
    alpha = 0.02 * float(np.abs(y).max()) / float(r0.abs().max())
    c = _dev(torch, rng.normal(size=y.shape))

    def step(z, th):
        return z - alpha * A.interior_residual(form_mu(float(th)), z, th, act, xt)

    loss = lambda z: (c * z).sum()
    th = torch.tensor(0.04, dtype=torch.float64, device="cuda")
    y0 = _dev(torch, y)
    L2, gy2, gt2 = A.checkpointed_adjoint(step, y0, th, 4, 2, loss)
    L4, gy4, gt4 = A.checkpointed_adjoint(step, y0, th, 4, 4, loss)
    assert L2 == L4
    assert torch.allclose(gy2, gy4, rtol=1e-13, atol=0) and torch.allclose(gt2, gt4, rtol=1e-13, atol=0)

    def run(yy, mu):
        z = yy
        with torch.no_grad():
            for _ in range(4):
                z = step(z, torch.tensor(mu, dtype=torch.float64, device="cuda"))
        return float(loss(z))

    d = _dev(torch, rng.normal(size=y.shape))
    h = 1e-6 * float(y0.abs().max()) / float(d.abs().max())
    fd = (run(y0 + h * d, 0.04) - run(y0 - h * d, 0.04)) / (2 * h)
    an = float((gy2 * d).sum())
    print("y: fd", fd, "adjoint", an)
    assert abs(fd - an) <= 1e-6 * abs(fd)
    hm = 1e-6 * 0.04
    fdm = (run(y0, 0.04 + hm) - run(y0, 0.04 - hm)) / (2 * hm)
    print("mu: fd", fdm, "adjoint", float(gt2))
    assert abs(fdm - float(gt2)) <= 1e-6 * abs(fdm)


@pytest.mark.parametrize("codes", ["reference", "random"])
def test_vjp_with_boundary_elements(env, oracle_lib, codes):
    """The VJP of the residual with the boundary rows (crimson_form_set_boundary)
    against central differences of the oracle's interior + boundary residual,
    for y and mu; back-flow at part of the outlet exercises the influx term,
    random codes every bit (prescribed values carry no derivative)."""
    cfa, torch = env
    from crimsonflowsolver_amd import mesh as M
    m, y, ac, rng = _state((5, 5, 7), 47)
    back = m.outlet & (m.x[:, 0] > 0.0)
    y[back, 2] = -np.abs(y[back, 2]) - 2.0
    ienb, iBCB, BCB = M.cylinder_boundary(m)
    if codes == "random":
        nelb = ienb.shape[0]
        pick = np.array([0, 1, 2, 4, 16, 3, 5, 6, 7], np.int32)
        iBCB = np.asfortranarray(np.stack([pick[rng.integers(0, pick.size, nelb)], iBCB[:, 1]], 1).astype(np.int32))
        BCB = np.asfortranarray(rng.uniform(-3, 3, (nelb, 3, 5)))
    colm, rowp = oracle_lib.genadj(m.ien, m.numnp)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp)
    form.set_boundary(ienb, iBCB, BCB)
    w = np.asfortranarray(rng.normal(size=(m.numnp, 4)))
    ybar, mub = _vjp(env, form, m, y, ac, w)

    def J(yy, mu=0.04):
        r, _, _ = oracle_lib.elmgmr_interior(m.ien, yy, ac, m.x, colm, rowp, oracle_lib.make_params(mu=mu, lhs=0))
        r = oracle_lib.boundary_res(ienb, iBCB, BCB, yy, m.x, r, mu=mu)
        return float(np.sum(w * r))

    for k in range(3):
        d = np.asfortranarray(rng.normal(size=y.shape) * np.array([1.0, 1.0, 1.0, 30.0]))
        h = 1e-6 * np.abs(y).max() / np.abs(d).max()
        fd = (J(y + h * d) - J(y - h * d)) / (2 * h)
        an = float(np.sum(ybar * d))
        print(codes, "direction", k, "fd", fd, "vjp", an)
        assert abs(fd - an) <= 1e-6 * max(abs(fd), 1e-300)
    hm = 1e-6 * 0.04
    fdm = (J(y, 0.04 + hm) - J(y, 0.04 - hm)) / (2 * hm)
    print(codes, "mu: fd", fdm, "vjp", mub)
    assert abs(fdm - mub) <= 1e-6 * abs(fdm)
