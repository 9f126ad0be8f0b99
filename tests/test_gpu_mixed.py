"""Mixed-precision form on the GPU, through the C ABI: the FP64 residual with
the tangent (iconvflow = 2: tauBar and u-bar terms, so the strong residual
is formed too) formed and assembled in FP32 (precision = 32), and memLS
solves on an FP32 matrix with FP32 Krylov arithmetic.

Tolerances, normwise relative (max|a - b| / max|b|):
  residual                  1e-12 against the oracle (it is the FP64 path);
  FP32 tangent              3e-6 against the oracle's FP64 tangent: every
                            entry is a sum of up to 24 element terms rounded
                            at FP32 epsilon 6e-8, with cancellation between
                            the mass, convection and viscous parts;
  FP32 solve                1e-3 against the direct FP64 solution at NS
                            relTol 1e-4 (the FP64 test bar is 10 x relTol).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def cfa():
    import crimsonflowsolver_amd as cfa
    if cfa.device_count() < 1:
        pytest.fail("no GPU visible: the HIP path must run here")
    return cfa


def _case(grid, seed):
    from crimsonflowsolver_amd import mesh as M
    m = M.cylinder_mesh(*grid)
    y, ac = M.poiseuille_state(m, seed=seed)
    rng = np.random.default_rng(seed)
    ac[:, :3] = rng.uniform(-50, 50, size=(m.numnp, 3))
    return m, y, ac


@pytest.mark.parametrize("grid,seed", [((6, 6, 10), 31), ((10, 10, 24), 32)])
def test_fp32_tangent_against_oracle(cfa, oracle_lib, grid, seed):
    m, y, ac = _case(grid, seed)
    colm, rowp = oracle_lib.genadj(m.ien, m.numnp)
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, params=cfa.form_params(precision=32, aspect_fp64=1e9))
    res, K, P = form(y, ac, m.x)
    rres, rK, rP = oracle_lib.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, oracle_lib.make_params())
    errs = {"res": _rel(res, rres), "lhsK": _rel(K, rK), "lhsP": _rel(P, rP)}
    print(grid, seed, errs)
    assert errs["res"] <= 1e-12
    assert errs["lhsK"] <= 3e-6 and errs["lhsP"] <= 3e-6
    assert errs["lhsK"] > 1e-10  # the tangent really was formed in FP32


def test_fp64_tau_island(cfa, oracle_lib):
    """aspect_fp64 = 1 puts every element in the FP64 island (metric, tau and
    the element record in FP64, rounded once to FP32), 1e9 none.  On a mesh
    compressed 50x along x (aspect ratios 50-150) the island's tangent
    stays within the FP32 bar; both errors are printed."""
    m, y, ac = _case((6, 6, 10), 33)
    m.x[:, 0] *= 0.02
    colm, rowp = oracle_lib.genadj(m.ien, m.numnp)
    rres, rK, rP = oracle_lib.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, oracle_lib.make_params())
    errs = {}
    for thr in (1.0, 1e9):
        form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, params=cfa.form_params(precision=32, aspect_fp64=thr))
        res, K, P = form(y, ac, m.x)
        errs[thr] = (_rel(K, rK), _rel(P, rP))
    print("island (K, P) err", errs[1.0], "all-FP32 (K, P) err", errs[1e9])
    assert errs[1.0][0] <= 3e-6 and errs[1.0][1] <= 3e-6


def test_mixed_fused_matrix_and_solve(cfa, oracle_lib):
    """precision-32 formation into the memLS matrix leaves an FP32 matrix
    whose blocks match the FP64 formation's; Res4 is the FP64 residual; the
    FP32 NS solve agrees with the direct FP64 solution."""
    import torch
    from crimsonflowsolver_amd import mesh as M
    m = M.cylinder_mesh(6, 6, 10)
    y, ac = M.poiseuille_state(m, seed=34)
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    ibc, bc = M.dirichlet_codes(m)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
    out = {}
    for prec in (64, 32):
        form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, iBC=ibc, params=cfa.form_params(precision=prec))
        lhs = cfa.memLS_LHS_CREATE(m.numnp, m.numnp, rowp.size, None, colm, rowp, 1)
        fixed = np.nonzero(ibc)[0] + 1
        cfa.memLS_BC_CREATE(lhs, 1, fixed.size, 3, cfa.BC_TYPE_Dir, fixed, np.zeros((3, fixed.size), order="F"))
        r4 = torch.zeros((m.numnp, 4), dtype=torch.float64, device="cuda")
        form.to_lhs(T(y), T(ac), T(m.x), r4, lhs)
        form.bc3res(r4, res4_layout=True)
        val = lhs.get_val(4)
        rhs = r4.cpu().numpy().T.copy(order="F")
        ls = cfa.memLS_LS_CREATE(cfa.LS_TYPE_NS, relTol=1e-4, dimKry=40, maxItr=10,
                                 relTolIn=[1e-4, 1e-4], maxItrIn=[10, 500])
        cfa.memLS_SOLVE_dev(lhs, ls, 4, r4)
        out[prec] = (val, rhs, r4.cpu().numpy().T.copy(), ls.RI)
    v64, b64, x64, s64 = out[64]
    v32, b32, x32, s32 = out[32]
    print("val", _rel(v32, v64), "res4", _rel(b32, b64), "stats", s64, s32)
    assert _rel(b32, b64) <= 1e-12
    assert _rel(v32, v64) <= 3e-6
    assert s32["suc"] == 1
    # direct FP64 solve of the FP64 system, scaled as PRECOND does
    from test_gpu_memls import _scaled_direct, _fixed_dofs
    xd = _scaled_direct(colm, rowp, v64, b64, 4, _fixed_dofs(ibc, 4))
    scale = np.abs(xd).max()
    print("x32 err", np.abs(x32 - xd).max() / scale, "x64 err", np.abs(x64 - xd).max() / scale)
    assert np.abs(x32 - xd).max() <= 1e-3 * scale


def test_precision_switch_roundtrip(cfa, oracle_lib):
    """set_precision(32) rounds the held planes to FP32; back to 64 keeps
    those rounded values; get_val reads either."""
    m, y, ac = _case((5, 5, 6), 35)
    colm, rowp = oracle_lib.genadj(m.ien, m.numnp)
    res, K, P = oracle_lib.elmgmr_interior(m.ien, y, ac, m.x, colm, rowp, oracle_lib.make_params())
    val = oracle_lib.pack_val4(colm, rowp, K, P)
    lhs = cfa.memLS_LHS_CREATE(m.numnp, m.numnp, rowp.size, None, colm, rowp, 0)
    lhs.set_val(4, val)
    assert np.array_equal(lhs.get_val(4), val)
    lhs.set_precision(32)
    v32 = lhs.get_val(4)
    assert np.array_equal(v32, val.astype(np.float32).astype(np.float64))
    lhs.set_precision(64)
    assert np.array_equal(lhs.get_val(4), v32)
