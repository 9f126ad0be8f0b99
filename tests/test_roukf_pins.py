"""Pins of the ROUKF restatement (crimsonflowsolver_amd.roukf) by what the
mathematics fixes:

* the simplex sigma points: I 1 = 0 and alpha I I^T = 1 (p = 1..6);
* for a model whose state depends linearly on the reduced parameters and a
  linear observation with Gaussian error, each ROUKF step is the Kalman
  filter's update of the parameters (mean and covariance, closed form,
  computed here from the normal equations), over several sequential
  observations, whatever the simplex; the state's other entries follow.
The Verdandi simplex construction itself is not reproduced bit for bit:
results of a nonlinear model depend on the choice of simplex at second
order (parity unpinned there).
"""
import numpy as np
import pytest

from crimsonflowsolver_amd import roukf


@pytest.mark.parametrize("p", [1, 2, 3, 6])
def test_simplex_points(p):
    I = roukf.simplex_points(p)
    a = 1.0 / (p + 1)
    assert I.shape == (p, p + 1)
    assert np.allclose(I.sum(axis=1), 0.0, atol=1e-14)
    assert np.allclose(a * I @ I.T, np.eye(p), atol=1e-14)


@pytest.mark.parametrize("p,m", [(1, 1), (2, 3), (3, 2)])
def test_linear_parameter_estimation_is_the_kalman_filter(p, m):
    rng = np.random.default_rng(10 * p + m)
    nf = 7                                   # field entries, linear in theta
    A = rng.standard_normal((nf, p))
    theta_true = rng.standard_normal(p)
    theta0 = theta_true + rng.standard_normal(p)
    P = np.diag(rng.uniform(0.5, 2.0, p))
    x0 = np.concatenate([A @ theta0, theta0])
    L0 = np.vstack([A, np.eye(p)])           # d x / d theta
    W = np.diag(rng.uniform(0.1, 0.3, m))
    f = roukf.ROUKF(x0, L0, P, W)
    theta, Pk = theta0.copy(), P.copy()
    for step in range(4):
        Hf = rng.standard_normal((m, nf))     # observation of the fields
        y = Hf @ (A @ theta_true) + rng.standard_normal(m) * 0.1

        def advance(X):                       # the fields follow theta
            return np.vstack([A @ X[nf:], X[nf:]])

        f.forward(advance)
        f.analyze(lambda X: y[:, None] - Hf @ X[:nf])
        # Kalman filter on theta: y = (Hf A) theta + e
        H = Hf @ A
        S = H @ Pk @ H.T + W
        K = Pk @ H.T @ np.linalg.inv(S)
        theta = theta + K @ (y - H @ theta)
        Pk = Pk - K @ H @ Pk
        assert np.allclose(f.x[nf:], theta, rtol=1e-10, atol=1e-12), step
        assert np.allclose(f.x[:nf], A @ theta, rtol=1e-10, atol=1e-12), step
        assert np.allclose(f.covariance()[nf:, nf:], Pk, rtol=1e-9, atol=1e-12), step
