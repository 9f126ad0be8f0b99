"""Pins of the oracle's numpy step glue (itrPredict, itrYAlpha, itrCorrect,
itrBC, ElmpvsQ's resistance term) by what the generalized-alpha scheme and
the geometry fix."""
import numpy as np

from oracle import oracle as O
from crimsonflowsolver_amd import mesh as M


def _fields(n, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.uniform(-1, 1, (n, 4))), np.asfortranarray(rng.uniform(-1, 1, (n, 4)))


def test_predict_then_correct_is_newmark():
    """y_{n+1} = y_n + dt ((1 - gamma) a_n + gamma a_{n+1}) for the velocity
    when the predictor's acceleration is corrected to a_{n+1}."""
    p = O.make_params()
    yold, acold = _fields(50, 1)
    anew = np.asfortranarray(np.random.default_rng(2).uniform(-1, 1, (50, 4)))
    y, ac = O.itr_predict(yold, acold, p.gami)
    assert np.array_equal(y, yold)
    d = np.zeros((4, 50), order="F")
    d[:3] = (anew[:, :3] - ac[:, :3]).T
    y1, ac1 = O.itr_correct(y, ac, d, p.gami, p.alfi, p.Delt)
    exp = yold[:, :3] + p.Delt * ((1 - p.gami) * acold[:, :3] + p.gami * anew[:, :3])
    assert np.allclose(y1[:, :3], exp, rtol=0, atol=1e-15)
    assert np.allclose(ac1[:, :3], anew[:, :3], rtol=0, atol=1e-15)
    d[3] = 2.0
    y2, _ = O.itr_correct(y, ac, d, p.gami, p.alfi, p.Delt)
    assert np.allclose(y2[:, 3], yold[:, 3] + 2.0 * p.gami * p.alfi * p.Delt, rtol=0, atol=1e-15)


def test_yalpha_endpoints():
    yold, acold = _fields(30, 3)
    y, ac = _fields(30, 4)
    ya, aca = O.itr_yalpha(yold, acold, y, ac, 1.0, 1.0)
    assert np.array_equal(ya, y) and np.array_equal(aca[:, :3], ac[:, :3]) and not aca[:, 3].any()
    ya, aca = O.itr_yalpha(yold, acold, y, ac, 0.0, 0.0)
    assert np.array_equal(ya[:, :3], yold[:, :3]) and np.array_equal(ya[:, 3], y[:, 3])
    assert np.array_equal(aca[:, :3], acold[:, :3])


def test_itr_bc_codes():
    y, _ = _fields(6, 5)
    bc = np.asfortranarray(np.arange(42, dtype=float).reshape(6, 7))
    ibc = np.array([0, 56, 4, 60, 0, 0], np.int32)  # none, velocity, pressure, both
    out = O.itr_bc(ibc, bc, y)
    assert np.array_equal(out[0], y[0]) and np.array_equal(out[4], y[4]) and np.array_equal(out[5], y[5])
    assert np.array_equal(out[1, :3], bc[1, 2:5]) and out[1, 3] == y[1, 3]
    assert out[2, 3] == bc[2, 0] and np.array_equal(out[2, :3], y[2, :3])
    assert np.array_equal(out[3, :3], bc[3, 2:5]) and out[3, 3] == bc[3, 0]


def test_itr_bc_partial_codes_lie_on_the_bc3lhs_constraint():
    """Codes 1-6: the free components are kept and the constrained ones are
    set on the affine constraint whose linear part is the substitution that
    bc3lhs / bc3Res project with (u_c = offset_c + (P u_free)_c), the
    offsets being BC(:,3) and, for the second component, BC(:,5)."""
    from test_bc3lhs_pins import CONSTRAINED, subst
    rng = np.random.default_rng(4)
    for code in range(1, 7):
        y = np.asfortranarray(rng.standard_normal((3, 4)))
        bc = np.asfortranarray(rng.uniform(-2, 2, (3, 7)))
        ibc = np.array([0, code << 3, 0], np.int32)
        out = O.itr_bc(ibc, bc, y)
        cons = CONSTRAINED[code]
        free = [k for k in range(3) if k not in cons]
        assert np.array_equal(out[1, free], y[1, free]) and out[1, 3] == y[1, 3]
        u = np.zeros(3)
        u[free] = y[1, free]
        lin = subst(code, bc[1]) @ u
        offs = [bc[1, 2], bc[1, 4]]
        for i, c in enumerate(cons):
            assert abs(out[1, c] - (offs[i] + lin[c])) <= 1e-14, (code, c)
        assert np.array_equal(out[[0, 2]], y[[0, 2]])


def test_resistance_flow_of_uniform_velocity():
    """Q = U x outlet area for u = (0, 0, U); the term adds -R Q NABI."""
    m = M.cylinder_mesh(6, 6, 5)
    out = np.nonzero(m.outlet)[0]
    nabi = m.nabi()[out]
    y = np.zeros((m.numnp, 4), order="F")
    y[:, 2] = 3.0
    tris = m.outlet_tris
    area = 0.5 * np.linalg.norm(np.cross(m.x[tris[:, 1]] - m.x[tris[:, 0]], m.x[tris[:, 2]] - m.x[tris[:, 0]]),
                                axis=1).sum()
    res0 = np.zeros((m.numnp, 4), order="F")
    res, q = O.elmpvsq_resistance(res0, y, out, nabi, 1333.0)
    assert abs(q - 3.0 * area) <= 1e-12 * area
    assert np.allclose(res[out, :3], -1333.0 * q * nabi, rtol=1e-15, atol=0)
    assert not res[:, 3].any()
