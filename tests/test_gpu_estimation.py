"""The estimation client (crimsonflowsolver_amd.estimation): ROUKF over a
cylinder flow whose sigma points run as one ensemble on the GPU.  A truth
run with known viscosity and outlet resistance produces pressures at nodes
near the inlet; the filter, started off in the parameter, must move its
estimate most of the way to the truth within a few steps (the filter's
algebra itself is pinned against the Kalman filter in test_roukf_pins.py).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfa():
    import crimsonflowsolver_amd as cfa
    if cfa.device_count() < 1:
        pytest.fail("no GPU visible: the HIP path must run here")
    return cfa


def _truth(mesh, y0, ac0, mu, R, nsteps, obs, inflow):
    from crimsonflowsolver_amd.estimation import CylinderFlow
    flow = CylinderFlow(mesh, 1, inflow)
    Y, AC = y0[None], ac0[None]
    out = []
    for _ in range(nsteps):
        Y, AC = flow.advance(Y, AC, np.array([mu]), np.array([R]))
        out.append(Y[0][obs, 3].copy())
    return out


@pytest.mark.parametrize("names,theta0,truth", [(("mu",), [0.06], [0.04]), (("R",), [1500.0], [1000.0])])
def test_roukf_recovers_a_parameter(cfa, names, theta0, truth):
    from crimsonflowsolver_amd import mesh as M
    from crimsonflowsolver_amd.estimation import Estimator
    mesh = M.cylinder_mesh(5, 5, 12)
    y0, ac0 = M.poiseuille_state(mesh, seed=3)
    inflow = 10.0
    z = mesh.x[:, 2]
    obs = np.nonzero((z < z.min() + 0.25 * (z.max() - z.min())) & ~mesh.wall & ~mesh.inlet)[0][:12]
    mu_t = truth[0] if names == ("mu",) else 0.04
    R_t = truth[0] if names == ("R",) else 1000.0
    nsteps = 4
    ys = _truth(mesh, y0, ac0, mu_t, R_t, nsteps, obs, inflow)
    fixed = {"mu": 0.04, "R": 1000.0}
    P0 = [[(0.5 * (theta0[0] - truth[0])) ** 2 * 4]]
    est = Estimator(mesh, names, theta0, P0, fixed, y0, ac0, obs, 1.0, inflow)
    hist = [theta0[0]]
    for n in range(nsteps):
        hist.append(float(est.step(ys[n])[0]))
    print(names, "estimates", hist, "truth", truth[0])
    assert abs(hist[-1] - truth[0]) <= 0.2 * abs(theta0[0] - truth[0])
