"""Pins of the oracle's bc3lhs / bc3Res for 1- and 2-component velocity
codes (bc3lhs.f90:41-690, bc3res.f90:34-80) against what the mathematics
fixes, independent of the reference's entry-by-entry operations:

* a partial code constrains components c in terms of the free ones,
  u_c = -sum_f BC(c,f) u_f (BC(:,4:6) of the node), i.e. u = P u_free with
  the 3x3 substitution matrix P below (free columns only);
* the reference's column operations on block (i, node) are B -> B P, its row
  operations on block (node, j) are B -> P^T B, the constrained diagonal
  entries become one; bc3Res is r -> P^T r.  For codes 1-5 these are the
  projection S^T M S embedded in the 3x3 block; for code 6 the column and
  residual operations are, but the row operations the reference writes
  (row 3 from row 1, bc3lhs.f90:648-680) are not: parity unpinned for the
  code-6 rows (restated as written);
* the formation applies bc3lhs to each element before the assembly; on the
  device it runs after the assembly on the summed blocks, which is the same
  by linearity (the constrained diagonal entry summing to the number of
  elements, interior and boundary, holding the node).

Blocks: entry m = 1..9 of lhsK / xKebe is B[r, c] with m - 1 = 3 r + c
(the reference's drawing 1 2 3 / 4 5 6 / 7 8 9), r the row node's
component and c the column node's.
"""
import numpy as np
import pytest

CONSTRAINED = {1: (0,), 2: (1,), 4: (2,), 3: (0, 1), 5: (0, 2), 6: (1, 2)}


def subst(code, bc):
    """u = P u_free for the node's code and BC(4:6) = bc[3:6]."""
    b4, b5, b6 = bc[3], bc[4], bc[5]
    P = np.eye(3)
    if code == 1:
        P[0] = [0, -b4, -b5]
    elif code == 2:
        P[1] = [-b4, 0, -b5]
    elif code == 4:
        P[2] = [-b4, -b5, 0]
    elif code == 3:
        P = np.array([[0, 0, -b4], [0, 0, -b6], [0, 0, 1.0]])
    elif code == 5:
        P = np.array([[0, -b4, 0], [0, 1.0, 0], [0, -b6, 0]])
    elif code == 6:
        P = np.array([[1.0, 0, 0], [-b4, 0, 0], [-b6, 0, 0]])
    return P


def _blocks(xK):
    return {(a, b): xK[:, a, b].reshape(3, 3).copy() for a in range(4) for b in range(4)}


def _expected(xK, codes, bc):
    """Per-element operations written as matrix products: columns, rows,
    then the ones (codes 0 and 7 untouched)."""
    B = _blocks(xK)
    P = {a: subst(codes[a], bc[a]) if codes[a] in CONSTRAINED else None for a in range(4)}
    for (a, b), M in B.items():
        if P[b] is not None:
            M = M @ P[b]
        if P[a] is not None:
            M = P[a].T @ M
        if a == b and P[a] is not None:
            for c in CONSTRAINED[codes[a]]:
                M[c, c] = 1.0
        B[(a, b)] = M
    out = np.zeros_like(xK)
    for (a, b), M in B.items():
        out[:, a, b] = M.reshape(9)
    return out


def _elem_case(rng, codes):
    ibc = np.array([c << 3 for c in codes], np.int32)
    bc = rng.uniform(-2, 2, (4, 7))
    xK = np.asfortranarray(rng.standard_normal((9, 4, 4)))
    return ibc, bc, xK


@pytest.mark.parametrize("code", [1, 2, 3, 4, 5])
def test_single_partial_node_is_the_projection(oracle_lib, code):
    rng = np.random.default_rng(100 + code)
    for inod in range(4):
        codes = [0, 0, 0, 0]
        codes[inod] = code
        ibc, bc, xK = _elem_case(rng, codes)
        got = oracle_lib.bc3lhs_elem(ibc, bc, np.arange(1, 5), xK)
        want = _expected(xK, codes, bc)
        assert np.allclose(got, want, rtol=0, atol=1e-13), (code, inod)
        # the constrained rows and columns are gone, the free block is S^T M S
        S = subst(code, bc[inod])[:, [f for f in range(3) if f not in CONSTRAINED[code]]]
        D = got[:, inod, inod].reshape(3, 3)
        free = [f for f in range(3) if f not in CONSTRAINED[code]]
        assert np.allclose(D[np.ix_(free, free)], S.T @ xK[:, inod, inod].reshape(3, 3) @ S, atol=1e-13)


def test_several_partial_nodes_commute(oracle_lib):
    """Codes 0-5 and 7 on all four nodes: the local node order of the
    reference's passes does not matter (rows and columns commute)."""
    rng = np.random.default_rng(7)
    for _ in range(40):
        codes = list(rng.choice([0, 1, 2, 3, 4, 5, 7], 4))
        ibc, bc, xK = _elem_case(rng, codes)
        got = oracle_lib.bc3lhs_elem(ibc, bc, np.arange(1, 5), xK)
        assert np.allclose(got, _expected(xK, codes, bc), rtol=0, atol=1e-12), codes
        # and the element's node numbering (the pass order) does not either
        perm = rng.permutation(4)
        ibc_p = np.zeros(4, np.int32)
        ibc_p[perm] = ibc
        bc_p = np.zeros_like(bc)
        bc_p[perm] = bc
        got_p = oracle_lib.bc3lhs_elem(ibc_p, bc_p, perm + 1, xK)
        assert np.allclose(got_p, got, rtol=0, atol=1e-12)


def test_code6_columns_are_the_substitution_rows_are_not(oracle_lib):
    """Code 6: its column operations are B -> B P; its row operations as the
    reference writes them differ from P^T B (parity of those unpinned)."""
    rng = np.random.default_rng(66)
    codes = [0, 6, 0, 0]
    ibc, bc, xK = _elem_case(rng, codes)
    got = oracle_lib.bc3lhs_elem(ibc, bc, np.arange(1, 5), xK)
    P = subst(6, bc[1])
    for a in (0, 2, 3):  # blocks (a, node): columns only
        assert np.allclose(got[:, a, 1].reshape(3, 3), xK[:, a, 1].reshape(3, 3) @ P, atol=1e-13)
    for b in (0, 2, 3):  # blocks (node, b): rows 1 and 2 zero, row 3 changed
        G = got[:, 1, b].reshape(3, 3)
        assert np.all(G[:2] == 0.0)
        assert not np.allclose(G, P.T @ xK[:, 1, b].reshape(3, 3))


@pytest.mark.parametrize("code", [1, 2, 3, 4, 5, 6])
def test_bc3res_partial_codes_are_the_projection(oracle_lib, code):
    rng = np.random.default_rng(200 + code)
    n = 5
    res = np.asfortranarray(rng.standard_normal((n, 4)))
    bc = np.asfortranarray(rng.uniform(-2, 2, (n, 7)))
    ibc = np.zeros(n, np.int32)
    ibc[2] = code << 3
    got = oracle_lib.bc3res(ibc, res, bc)
    want = res.copy()
    P = subst(code, bc[2])
    want[2, :3] = P.T @ res[2, :3]
    for c in CONSTRAINED[code]:
        want[2, c] = 0.0
    assert np.allclose(got, want, rtol=0, atol=1e-14)


def _mesh_with_partial_codes(seed):
    from crimsonflowsolver_amd import mesh as M
    m = M.cylinder_mesh(4, 4, 6)
    rng = np.random.default_rng(seed)
    codes = np.zeros(m.numnp, np.int32)
    wall = np.nonzero(m.wall)[0] if hasattr(m, "wall") else np.arange(m.numnp)
    pick = rng.choice(wall, size=min(40, wall.size), replace=False)
    codes[pick] = rng.choice([1, 2, 3, 4, 5], pick.size)
    # a few code-6 nodes sharing no element with another partial node
    ien0 = m.ien - 1
    for cand in rng.permutation(m.numnp):
        if codes[cand]:
            continue
        elems = np.nonzero((ien0 == cand).any(axis=1))[0]
        if not (codes[ien0[elems]] != 0).any():
            codes[cand] = 6
            if (codes == 6).sum() >= 3:
                break
    bc = np.asfortranarray(rng.uniform(-1.5, 1.5, (m.numnp, 7)))
    return m, (codes << 3).astype(np.int32), bc


def test_interior_bc3lhs_equals_the_post_assembly_form(oracle_lib):
    """The oracle's per-element bc3lhs summed over the elements equals the
    operations applied once to the assembled blocks (the device's order),
    with each constrained diagonal entry equal to the node's element count."""
    from crimsonflowsolver_amd import mesh as M
    m, ibc, bc = _mesh_with_partial_codes(3)
    y, ac = M.poiseuille_state(m, seed=3)
    colm, rowp = oracle_lib.genadj(m.ien, m.numnp)
    prm = oracle_lib.make_params()
    r0, K0, P0 = oracle_lib.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, prm)
    r1, K1, P1 = oracle_lib.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, prm, ibc=ibc, bc=bc)
    assert np.array_equal(r0, r1) and np.array_equal(P0, P1)  # bc3lhs touches xKebe only
    codes = ibc >> 3
    cnt = np.bincount(m.ien.ravel() - 1, minlength=m.numnp)
    want = K0.copy()
    for i in range(m.numnp):
        for k in range(colm[i] - 1, colm[i + 1] - 1):
            j = rowp[k] - 1
            if codes[i] not in CONSTRAINED and codes[j] not in CONSTRAINED:
                continue
            B = K0[:, k].reshape(3, 3)
            if codes[j] in CONSTRAINED:
                B = B @ subst(codes[j], bc[j])
            if codes[i] in CONSTRAINED:
                if codes[i] == 6:  # rows as the reference writes them (not P^T B)
                    G = np.zeros((3, 3))
                    G[2] = B[2] - bc[i, 3] * B[0]
                    G[2, 2] -= bc[i, 5] * B[1, 2]
                    B = G
                else:
                    B = subst(codes[i], bc[i]).T @ B
                if i == j:
                    for c in CONSTRAINED[codes[i]]:
                        B[c, c] = cnt[i]
            want[:, k] = B.reshape(9)
    scale = np.abs(K0).max()
    assert np.abs(K1 - want).max() <= 1e-12 * scale
    assert (codes == 6).sum() >= 1 and (np.isin(codes, [1, 2, 3, 4, 5])).sum() >= 10


def test_boundary_loop_adds_one_per_boundary_element(oracle_lib):
    from crimsonflowsolver_amd import mesh as M
    m, ibc, bc = _mesh_with_partial_codes(5)
    ienb, _, _ = M.cylinder_boundary(m)
    colm, rowp = oracle_lib.genadj(m.ien, m.numnp)
    K = oracle_lib.boundary_lhs(ienb, ibc, bc, colm, rowp, np.zeros((9, rowp.size), order="F"))
    codes = ibc >> 3
    nb = np.bincount(np.asarray(ienb).ravel() - 1, minlength=m.numnp)
    for i in range(m.numnp):
        k = colm[i] - 1 + int(np.nonzero(rowp[colm[i] - 1:colm[i + 1] - 1] == i + 1)[0][0])
        D = K[:, k].reshape(3, 3)
        want = np.zeros((3, 3))
        if codes[i] in CONSTRAINED:
            for c in CONSTRAINED[codes[i]]:
                want[c, c] = nb[i]
        assert np.array_equal(D, want), i
    off = np.ones(rowp.size, bool)
    for i in range(m.numnp):
        off[colm[i] - 1 + int(np.nonzero(rowp[colm[i] - 1:colm[i + 1] - 1] == i + 1)[0][0])] = False
    assert np.all(K[:, off] == 0.0)
