"""Reduced-order unscented Kalman filter with simplex sigma points, the
estimator the reference runs its flow solver under (ROUKFModified,
estimation/src/ROUKFModified.cxx:105-260 Forward, :780-945 the analysis;
Verdandi's ReducedOrderUnscentedKalmanFilter; Moireau & Chapelle, ESAIM
COCV 17 (2011)).

The state x (flow fields and the estimated parameters, any array) carries
a reduced error space of dimension p: L (the state's sensitivity to the p
reduced directions) and U^{-1} (the reduced covariance).  One step:

  forward   C = chol(U^{-1}); x_i = x + L C I_i for the p + 1 simplex points
            I_i; each x_i is advanced by the model (one ensemble, formed and
            solved as one union on the GPU); x = alpha sum x_i
  analysis  L = alpha X I^T; z_i = y - H(x_i) (innovations), zbar = alpha
            sum z_i; HL^T = -alpha I Z^T; U = 1 + HL^T W^{-1} HL;
            x += L U^{-1} HL^T W^{-1} zbar

with alpha = 1/(p+1) and I (p x (p+1)) a simplex with I 1 = 0 and alpha I
I^T = 1.  For a model and observation linear in the reduced directions the
step is the Kalman filter's (its covariance L U^{-1} L^T).  The per-state
work (the forward runs) is the caller's; this module holds the p x p
algebra, on the host as in the reference.
"""
from __future__ import annotations

import numpy as np


def simplex_points(p):
    """I (p x (p+1)): row j has -1/sqrt(alpha j (j+1)) in its first j
    columns, j/sqrt(alpha j (j+1)) in column j, zeros after (j = 1..p), so
    that I 1 = 0 and alpha I I^T = 1 (alpha = 1/(p+1))."""
    alpha = 1.0 / (p + 1)
    I = np.zeros((p, p + 1))
    for j in range(1, p + 1):
        s = np.sqrt(alpha * j * (j + 1))
        I[j - 1, :j] = -1.0 / s
        I[j - 1, j] = j / s
    return I


class ROUKF:
    """x0: initial state (1-D numpy array, or anything supporting + and
    scalar *); L0: (n x p) initial sensitivities (numpy array of the state's
    flattened size); P0: (p x p) initial reduced covariance; W: observation
    error covariance (m x m)."""

    def __init__(self, x0, L0, P0, W):
        self.x = np.array(x0, dtype=np.float64)
        self.L = np.array(L0, dtype=np.float64)
        self.Uinv = np.array(P0, dtype=np.float64)
        self.Winv = np.linalg.inv(np.atleast_2d(np.array(W, dtype=np.float64)))
        self.p = self.L.shape[1]
        self.I = simplex_points(self.p)
        self.alpha = 1.0 / (self.p + 1)

    def sigma_points(self):
        """The p + 1 states x_i = x + L C I_i (columns of the result)."""
        C = np.linalg.cholesky(self.Uinv)
        return self.x[:, None] + (self.L @ C) @ self.I

    def forward(self, advance):
        """advance(X) -> the p + 1 states after one step (columns)."""
        self.X = np.asarray(advance(self.sigma_points()), dtype=np.float64)
        self.x = self.alpha * self.X.sum(axis=1)
        return self.x

    def analyze(self, innovations):
        """innovations(X) -> Z (m x (p+1)), z_i = y - H(x_i)."""
        Z = np.atleast_2d(np.asarray(innovations(self.X), dtype=np.float64))
        self.L = self.alpha * self.X @ self.I.T
        zbar = self.alpha * Z.sum(axis=1)
        HLt = -self.alpha * self.I @ Z.T          # (p x m) = (H L)^T
        wm = HLt @ self.Winv
        self.Uinv = np.linalg.inv(np.eye(self.p) + wm @ HLt.T)
        self.x = self.x + self.L @ (self.Uinv @ (wm @ zbar))
        return self.x

    def covariance(self):
        return self.L @ self.Uinv @ self.L.T
