"""PHASTA file reader/writer and the real-geometry fixture (the reference's
mainTests/basic cylinder, tests/golden/basic_case.npz).  CPU only."""
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(__file__), "golden", "basic_case.npz")


def load_case():
    z = np.load(GOLD)
    return {k: z[k] for k in z.files}


def test_restart_roundtrip(tmp_path):
    from crimsonflowsolver_amd import phasta
    rng = np.random.default_rng(1)
    y, ac = rng.normal(size=(37, 5)), rng.normal(size=(37, 5))
    p = tmp_path / "restart.12.1"
    phasta.write_restart(str(p), y, ac, 12)
    r = phasta.read_restart(str(p))
    assert r["step"] == 12
    assert np.array_equal(r["y"], y) and np.array_equal(r["ac"], ac)
    s = phasta.to_solver_order(r["y"])
    assert np.array_equal(s[:, :3], y[:, 1:4]) and np.array_equal(s[:, 3], y[:, 0])


def test_basic_case_fixture():
    c = load_case()
    x, ien = c["x"], c["ien"]
    assert x.shape == (1072, 3) and ien.shape == (4157, 4)
    assert ien.min() == 1 and ien.max() == 1072
    xe = x[ien - 1]
    V = np.einsum("ei,ei->e", xe[:, 0] - xe[:, 3], np.cross(xe[:, 1] - xe[:, 3], xe[:, 2] - xe[:, 3])) / 6
    assert (V > 0).all()
    # the tets fill the cylinder r = 3, -20 <= z <= 20 up to the faceting of
    # its wall (a coarse polygon: 2.6 % less area than the circle)
    assert 0.95 * np.pi * 9 * 40 < V.sum() < np.pi * 9 * 40
    assert set(np.unique(c["iBC"])) == {0, 56} and (c["iBC"] == 56).sum() == 26


def test_oracle_patch_test_on_real_geometry(oracle_lib):
    """Constant state on the real mesh: the assembled residual vanishes at
    every node off the boundary faces (the P1 patch test)."""
    c = load_case()
    x, ien = np.asfortranarray(c["x"]), np.asfortranarray(c["ien"])
    n = x.shape[0]
    colm, rowp = oracle_lib.genadj(ien, n)
    y = np.zeros((n, 4), order="F")
    y[:, :3] = [0.5, -1.0, 12.0]
    y[:, 3] = 1000.0
    res, _, _ = oracle_lib.elmgmr_interior(ien, y, np.zeros_like(y), x, colm, rowp, oracle_lib.make_params())
    bnd = np.unique(c["ienb"][:, :3].ravel() - 1)
    interior = np.setdiff1d(np.arange(n), bnd)
    assert interior.size > 100
    assert np.abs(res[interior]).max() < 1e-10 * np.abs(res).max()


def _write_geombc(path, x, ien_blocks, bnd_blocks):
    """A geombc file with several interior and boundary element blocks (the
    layout read_geombc reads; each boundary block: connectivity, then its
    nbc codes and nbc values, as PHASTA's writer orders them)."""
    from crimsonflowsolver_amd import phasta as P
    parts = [b"# PHASTA Input File Version 2.0\n",
             P._block("byteorder magic number", [1], np.array([P.MAGIC], "<i4").tobytes()),
             P._block("number of nodes", [x.shape[0]]),
             P._block("co-ordinates", [x.shape[0], 3], np.asfortranarray(x, "<f8").tobytes(order="F"))]
    for ien in ien_blocks:
        parts.append(P._block("connectivity interior linear tetrahedron", [ien.shape[0], 4, 0, 1, 4],
                              np.asfortranarray(ien, "<i4").tobytes(order="F")))
    for ienb, ibcb, bcb in bnd_blocks:
        n = ienb.shape[0]
        parts.append(P._block("connectivity boundary linear tetrahedron", [n, 4, 0, 1, 4, 3],
                              np.asfortranarray(ienb, "<i4").tobytes(order="F")))
        parts.append(P._block("nbc codes linear tetrahedron", [n, 2],
                              np.asfortranarray(ibcb, "<i4").tobytes(order="F")))
        parts.append(P._block("nbc values linear tetrahedron", [n, bcb.shape[2]],
                              np.asfortranarray(bcb, "<f8").tobytes(order="F")))
    with open(path, "wb") as f:
        f.write(b"".join(parts))


def test_geombc_with_several_element_blocks(tmp_path):
    """Every interior and boundary block is kept, concatenated in file order,
    with each boundary block's codes and values beside its elements."""
    from crimsonflowsolver_amd import phasta
    rng = np.random.default_rng(3)
    x = rng.normal(size=(20, 3))
    ien = [rng.integers(1, 21, size=(n, 4)).astype(np.int32) for n in (5, 3)]
    bnd = [(rng.integers(1, 21, size=(n, 4)).astype(np.int32), rng.integers(0, 9, size=(n, 2)).astype(np.int32),
            rng.normal(size=(n, 3, 2))) for n in (4, 2, 3)]
    p = tmp_path / "geombc.dat.1"
    _write_geombc(str(p), x, ien, bnd)
    g = phasta.read_geombc(str(p))
    assert np.array_equal(g["x"], x)
    assert np.array_equal(g["ien"], np.concatenate(ien))
    assert np.array_equal(g["ienb"], np.concatenate([b[0] for b in bnd]))
    assert np.array_equal(g["iBCB"], np.concatenate([b[1] for b in bnd]))
    assert np.array_equal(g["BCB"], np.concatenate([b[2] for b in bnd]))
    # a boundary block whose codes disagree with its elements is refused
    bad = [(bnd[0][0], bnd[1][1], bnd[0][2])]
    _write_geombc(str(p), x, ien, bad)
    import pytest
    with pytest.raises(ValueError):
        phasta.read_geombc(str(p))
