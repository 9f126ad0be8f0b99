"""The threaded build of the CPU restatement (oracle/liboracle_mt.so), timed
as the all-core CPU baseline, against the serial restatement.  CPU only.

  * the two-phase formation (cfo_elmgmr_interior_plan) gives the serial block
    loop's bits, serially and on several threads, with and without bc3lhs;
  * the threaded memLS NS solve reaches the serial solve's solution to the
    solver tolerance (its dot products add per-thread partial sums, so the
    iterates differ in the last bits).
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def case(oracle_lib):
    from crimsonflowsolver_amd import mesh as M
    O = oracle_lib
    m = M.cylinder_mesh(8, 8, 14)
    colm, rowp = O.genadj(m.ien, m.numnp)
    y, ac = M.poiseuille_state(m, seed=5)
    rng = np.random.default_rng(5)
    ac[:, :3] = rng.uniform(-50, 50, size=(m.numnp, 3))
    return O, M, m, colm, rowp, y, ac


@pytest.mark.parametrize("mt,threads", [(False, 1), (True, 1), (True, 3), (True, 8)])
def test_plan_formation_is_bitwise_the_block_loop(case, mt, threads):
    O, M, m, colm, rowp, y, ac = case
    p = O.make_params()
    want = O.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, p)
    if mt:
        O.lib_mt(threads)
    got = O.Plan(m.ien, m.numnp, mt=mt).interior(y, ac, m.x, colm, rowp, p)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    # residual only
    p0 = O.make_params(lhs=0)
    r, _, _ = O.Plan(m.ien, m.numnp, mt=mt).interior(y, ac, m.x, colm, rowp, p0)
    assert np.array_equal(r, O.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, p0)[0])


def test_plan_formation_with_partial_bc(case):
    """bc3lhs inside the element phase: codes 1-6 on some nodes."""
    O, M, m, colm, rowp, y, ac = case
    rng = np.random.default_rng(8)
    ibc = np.zeros(m.numnp, np.int32)
    sel = rng.choice(m.numnp, 40, replace=False)
    ibc[sel] = rng.integers(1, 7, size=40) << 3
    bc = np.zeros((m.numnp, 6), order="F")
    bc[:, 3:6] = rng.normal(size=(m.numnp, 3))
    p = O.make_params()
    want = O.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, p, ibc=ibc, bc=bc)
    O.lib_mt(4)
    got = O.Plan(m.ien, m.numnp, mt=True).interior(y, ac, m.x, colm, rowp, p, ibc=ibc, bc=bc)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


def test_threaded_ns_solve_matches_serial(case):
    O, M, m, colm, rowp, y, ac = case
    p = O.make_params()
    res, K, P = O.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, p)
    ibc, _ = M.dirichlet_codes(m)
    res = O.bc3res(ibc, res)
    fixed = np.nonzero(ibc)[0]
    sols = []
    for mt in (False, True):
        if mt:
            O.lib_mt(4)
        ls = O.MemLS(colm, rowp, 1, relTol=1e-6, dimKry=60, maxItr=10, relTolIn=[1e-3, 1e-3],
                     maxItrIn=[20, 400], mt=mt)
        ls.bc_create(1, fixed + 1, 3, O.BC_TYPE_Dir, np.zeros((3, fixed.size), order="F"))
        sols.append(ls.solve(np.asfortranarray(res.T), O.pack_val4(colm, rowp, K, P, mt=mt)))
        assert ls.stats(0)["suc"] == 1
    a, b = sols
    assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
    assert np.array_equal(O.pack_val4(colm, rowp, K, P, mt=True), O.pack_val4(colm, rowp, K, P))
