"""Outlet 0D models as implicit coefficients of the coupled surfaces.

ElmGMR couples a 0D model to a surface through two numbers per timestep
(ElmpvsQ, elmgmr.f90:1152-1330; AddElmpvsQFormemLS, :1775-1856): the
surface pressure at t_{n+alpha_f} as an affine function of the flow through
the surface, P = c1 Q + c2, with Q = sum NABI . u over the surface nodes at
y_alpha.  The residual gets -(c1 Q + c2) NABI on those nodes
(`ElmGMRInterior.elmpvsq_coef`, crimson_elmpvsq_coef_dev) and memLS's face
coupling coefficient is c1 alpha_f gamma Delt (`face_res`).

`Resistance` is c1 = R, c2 = 0.  `RCR` is the reference's convolution form
of the three-element Windkessel (P = Rp Q + Pc, C dPc/dt = Q - (Pc - Pd)/Rd):
exact for a flow linear between time steps, with the distal pressure's
initial-condition term and the alpha_f compensation for a flow taken from
y_alpha.  Coefficients are host scalars (a few per surface per step), as in
the reference; the per-node work runs on the device.
"""
from __future__ import annotations

import math

import numpy as np


class Resistance:
    """ElmpvsQ's resistance surfaces (elmgmr.f90:1152-1164)."""

    def __init__(self, R):
        self.R = float(R)

    def coefficients(self, step):
        return self.R, 0.0

    def advance(self, q_new):
        pass


class RCR:
    """One RCR surface (ValueListRCR(1:3) = Rp, C, Rd; distal pressure
    Pdist(t) from a periodic table (times, values), linear in between,
    bctint.f90:583-611).  `P0`, `Q0`: the surface's mean pressure and flow at
    the initial state (calcRCRic, bctcardio.f90:589-627).

    Step n (0-based: the step from t_n to t_{n+1}) uses
    `coefficients(n)` = (c1, c2) with
      c1 = beta_{n+2}                                 (RCRConvCoef(n+2))
      c2 = sum_{j <= n+1} QHist_j beta_j + H_n        (pHist + CalcHopRCR)
    and after the step `advance(Q_{n+1})` appends the flow at t_{n+1}
    (UpdHistConv, bctcardio.f90:121-123)."""

    def __init__(self, Rp, C, Rd, alfi, Delt, P0, Q0, pdist_times=(0.0, 1.0), pdist_values=(0.0, 0.0)):
        self.Rp, self.C, self.Rd = float(Rp), float(C), float(Rd)
        self.alfi, self.Delt = float(alfi), float(Delt)
        self.d = self.Delt / (self.C * self.Rd)  # dtRCR (genini.f90:126)
        self.pt = np.asarray(pdist_times, dtype=np.float64)
        self.pv = np.asarray(pdist_values, dtype=np.float64)
        if self.pt.size < 2 or self.pt.size != self.pv.size or np.any(np.diff(self.pt) <= 0):
            raise ValueError("pdist_times must rise and match pdist_values")
        # initial-condition term, at currentTimestepIndex = 0 (bctcardio.f90:613-621)
        self.ic = float(P0) - self.Rp * float(Q0) - self.pdist(0.0)
        self.qhist = [float(Q0)]

    def pdist(self, t):
        """RCRint (bctint.f90:593-609): periodic in the table's last time."""
        period = self.pt[-1]
        nper = int(t / period)
        pt = t - nper * period
        for j in range(1, self.pt.size):
            if self.pt[j] > pt:
                wr = (pt - self.pt[j - 1]) / (self.pt[j] - self.pt[j - 1])
                return float(self.pv[j - 1] * (1.0 - wr) + self.pv[j] * wr)
        return 0.0  # the table's last time itself: left unset by the reference

    def conv_coef(self, step):
        """CalcRCRConvCoef (bctcardio.f90:149-190), beta_1..beta_{n+2}
        (returned 0-based), compensated for a flow taken at y_alpha."""
        a, d, Rp, Rd = self.alfi, self.d, self.Rp, self.Rd
        n = int(step)
        b = np.zeros(n + 2)
        ea = math.exp(-a * d)
        if n == 0:
            b[0] = Rp * (1.0 - a) + Rd * (-a + 1.0 + 1.0 / d - ea * (1.0 + 1.0 / d))
            b[1] = Rp * a + Rd * (a - 1.0 / d + ea / d)
        else:
            b[0] = -Rd * math.exp(-d * (n + a)) * (1.0 + (1.0 - math.exp(d)) / d)
            b[n] = Rp * (1.0 - a) - Rd * (a - 1.0 - 1.0 / d + ea / d * (2.0 - math.exp(-d)))
            b[n + 1] = Rp * a + Rd * (a - 1.0 / d + ea / d)
            for j in range(2, n + 1):  # 1-based j = 2..n
                b[j - 1] = Rd / d * math.exp(-d * (n + a + 2 - j)) * (1.0 - math.exp(d)) ** 2
        b[n] = b[n] - b[n + 1] * (1.0 - a) / a
        b[n + 1] = b[n + 1] / a
        return b

    def coefficients(self, step):
        n = int(step)
        if len(self.qhist) != n + 1:
            raise ValueError(f"step {n} needs the flow history up to t_{n} ({len(self.qhist)} values held)")
        b = self.conv_coef(n)
        pold = 0.0
        for j in range(n + 1):  # pHist (bctint.f90:1141-1146)
            pold = pold + self.qhist[j] * b[j]
        hop = self.ic * math.exp(-self.d * (n + self.alfi)) + self.pdist(self.Delt * (n + self.alfi))
        return float(b[n + 1]), pold + hop

    def advance(self, q_new):
        self.qhist.append(float(q_new))


def face_res(c1, alfi, gami, Delt):
    """memLS face coupling coefficient of a surface (elmgmr.f90:1853)."""
    return float(c1) * alfi * gami * Delt
