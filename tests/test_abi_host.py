"""CPU checks of the C ABI (include/crimson_amd.h): the library loads, every
declared entry point is exported, and the host-side validation paths (which
run before any HIP call) return the documented status codes.  Also the
native genadj against the oracle and the mesh generator against its numpy
twin.  No GPU needed."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def cfa():
    import crimsonflowsolver_amd as cfa
    cfa.lib()
    return cfa


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "crimson_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(crimson_\w+)\s*\(", txt)))


def test_every_declared_symbol_is_exported(cfa):
    names = declared_functions()
    assert len(names) >= 25
    L = cfa.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(cfa.EXPORTS)


def test_device_count_without_gpu_is_safe(cfa):
    assert cfa.device_count() >= 0


def test_genadj_matches_oracle(cfa, oracle_lib):
    from crimsonflowsolver_amd import mesh as M
    for g in ((3, 3, 3), (6, 5, 7)):
        m = M.cylinder_mesh(*g)
        a = cfa.genadj(m.ien, m.numnp)
        b = oracle_lib.genadj(m.ien, m.numnp)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # nodes 5 and 7 belong to no element: their rows stay empty (Asadj adds
    # nothing for them, genadj.f90:29-44), and the formation rejects such a
    # graph (no diagonal block)
    ien = np.asfortranarray(np.array([[1, 2, 3, 4], [2, 3, 4, 6]], np.int32))
    a = cfa.genadj(ien, 7)
    b = oracle_lib.genadj(ien, 7)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert a[0][4] == a[0][5] and a[0][6] == a[0][7]
    with pytest.raises(cfa.CrimsonError) as e:
        cfa.ElmGMRInterior(ien, a[0], a[1], 7)
    assert e.value.code == 4 and "row 5" in str(e.value)


def test_genadj_capacity_error(cfa):
    from crimsonflowsolver_amd import mesh as M
    m = M.cylinder_mesh(3, 3, 3)
    ien = np.asfortranarray(m.ien)
    colm = np.zeros(m.numnp + 1, np.int32)
    rowp = np.zeros(5, np.int32)
    nnz = C.c_int()
    rc = cfa.lib().crimson_genadj(m.numnp, m.nel, ien.ctypes.data_as(C.c_void_p), colm.ctypes.data_as(C.c_void_p),
                                  rowp.ctypes.data_as(C.c_void_p), 5, C.byref(nnz))
    assert rc == 1 and nnz.value > 5


def test_mesh_generator_matches_numpy_twin():
    from crimsonflowsolver_amd import mesh as M
    for g in ((4, 4, 3), (7, 5, 4), (6, 6, 2)):
        a, b = M.cylinder_mesh(*g), M.cylinder_mesh_numpy(*g)
        assert np.array_equal(a.ien, b.ien) and np.array_equal(a.x, b.x)
        for k in ("inlet", "outlet", "wall"):
            assert np.array_equal(getattr(a, k), getattr(b, k))
        np.testing.assert_allclose(a.nabi(), b.nabi(), atol=1e-15)
    m = M.cylinder_mesh(8, 8, 4)
    x = m.x[m.ien - 1]
    V = np.einsum("ei,ei->e", x[:, 0] - x[:, 3], np.cross(x[:, 1] - x[:, 3], x[:, 2] - x[:, 3])) / 6.0
    assert (V > 0).all()
    # outlet area = kept columns x cell area
    assert m.nabi()[:, 2].sum() == pytest.approx(np.abs(m.nabi()[:, 2]).sum())


def test_ls_create_defaults_and_limits(cfa):
    ls = cfa.memLS_LS_CREATE(cfa.LS_TYPE_NS)
    assert ls.LS_type == cfa.LS_TYPE_NS
    with pytest.raises(cfa.CrimsonError) as e:
        cfa.memLS_LS_CREATE(7)
    assert e.value.code == 1
    with pytest.raises(cfa.CrimsonError) as e:
        cfa.memLS_LS_CREATE(cfa.LS_TYPE_NS, maxItr=200)
    assert e.value.code == 2
    cfa.memLS_LS_CREATE(cfa.LS_TYPE_CG, maxItr=5000)  # CG has no projection-count limit


def _small_graph():
    from crimsonflowsolver_amd import mesh as M
    m = M.cylinder_mesh(3, 3, 3)
    import crimsonflowsolver_amd as cfa
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    return m, colm, rowp


@pytest.mark.parametrize("field,value,code", [("itau", 1, 2), ("iconvflow", 1, 2), ("idiff", 2, 2),
                                              ("ipord", 2, 2), ("precision", 16, 2), ("rho", -1.0, 1),
                                              ("ideformwall", 2, 1)])
def test_form_setup_rejects_unsupported_flags(cfa, field, value, code):
    m, colm, rowp = _small_graph()
    prm = cfa.form_params()
    setattr(prm, field, value)
    with pytest.raises(cfa.CrimsonError) as e:
        cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, params=prm)
    assert e.value.code == code


def test_form_setup_rejects_periodic_and_order_dependent_code6(cfa):
    """Periodic nodes, and a code-6 node sharing an element with another
    1- or 2-component node (the reference's code-6 row operations then depend
    on the element's local node order), are rejected before any GPU work."""
    m, colm, rowp = _small_graph()
    ibc = np.zeros(m.numnp, np.int32)
    ibc[3] = 1 << 10
    with pytest.raises(cfa.CrimsonError) as e:
        cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, iBC=ibc)
    assert e.value.code == 2
    for other in (8, 16, 8 | 32, 16 | 32):
        ibc = np.zeros(m.numnp, np.int32)
        ibc[m.ien[0, 0] - 1] = 16 | 32
        ibc[m.ien[0, 2] - 1] = other
        with pytest.raises(cfa.CrimsonError) as e:
            cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, iBC=ibc)
        assert e.value.code == 2 and "code 6" in str(e.value)


def test_form_setup_rejects_bad_graph_and_connectivity(cfa):
    m, colm, rowp = _small_graph()
    bad = m.ien.copy(order="F")
    bad[0, 0] = m.numnp + 5
    with pytest.raises(cfa.CrimsonError) as e:
        cfa.ElmGMRInterior(bad, colm, rowp, m.numnp)
    assert e.value.code == 1
    # drop row 1's diagonal: rowp loses an entry, colm shifts
    k = colm[0] - 1 + int(np.nonzero(rowp[colm[0] - 1:colm[1] - 1] == 1)[0][0])
    rowp2 = np.delete(rowp, k)
    colm2 = colm.copy()
    colm2[1:] -= 1
    with pytest.raises(cfa.CrimsonError) as e:
        cfa.ElmGMRInterior(m.ien, colm2, rowp2, m.numnp)
    assert e.value.code == 4
    # an element pair missing from the graph
    i = m.ien[0, 0] - 1
    j = m.ien[0, 1]
    row = rowp[colm[i] - 1:colm[i + 1] - 1]
    k = colm[i] - 1 + int(np.nonzero(row == j)[0][0])
    rowp3 = np.delete(rowp, k)
    colm3 = colm.copy()
    colm3[i + 1:] -= 1
    with pytest.raises(cfa.CrimsonError) as e:
        cfa.ElmGMRInterior(m.ien, colm3, rowp3, m.numnp)
    assert e.value.code in (3,)


def test_lhs_create_validation(cfa):
    m, colm, rowp = _small_graph()
    n = m.numnp
    with pytest.raises(cfa.CrimsonError) as e:
        cfa.memLS_LHS_CREATE(n + 1, n, rowp.size, None, colm, rowp, 1)
    assert e.value.code == 1
    g = np.arange(1, n + 1, dtype=np.int32)[::-1].copy()
    with pytest.raises(cfa.CrimsonError) as e:
        cfa.memLS_LHS_CREATE(n, n, rowp.size, g, colm, rowp, 1)
    assert e.value.code == 2
    bad = rowp.copy()
    bad[0] = n + 3
    with pytest.raises(cfa.CrimsonError) as e:
        cfa.memLS_LHS_CREATE(n, n, rowp.size, None, colm, bad, 1)
    assert e.value.code == 1


def test_no_cpu_fallback_in_product_path():
    """The product package never imports the oracle."""
    pkg = os.path.join(ROOT, "crimsonflowsolver_amd")
    for fn in os.listdir(pkg):
        if fn.endswith(".py"):
            assert "oracle" not in open(os.path.join(pkg, fn)).read().replace("oracle)", ""), fn


def test_unstructured_mesh_exercises_long_rows(cfa):
    """The jittered Delaunay mesh of the GPU parity test has rows longer than
    one 16-block group and nodes with more than two 16-incidence chunks."""
    from crimsonflowsolver_amd import mesh as M
    m = M.jittered_delaunay_mesh(7, 7, 12, seed=17)
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    assert np.diff(colm).max() > 16
    assert np.bincount(m.ien.ravel()).max() > 32
    x = m.x[m.ien - 1]
    V = np.einsum("ei,ei->e", x[:, 0] - x[:, 3], np.cross(x[:, 1] - x[:, 3], x[:, 2] - x[:, 3]))
    assert (V > 0).all()


def test_header_compiles_as_c99_and_cxx():
    """include/crimson_amd.h is usable from C (the Fortran host's ISO_C_BINDING
    interfaces mirror it) and from C++."""
    import shutil
    import subprocess
    hdr = os.path.join(ROOT, "include", "crimson_amd.h")
    for comp, std, lang in (("gcc", "-std=c99", "c"), ("g++", "-std=c++17", "c++")):
        exe = shutil.which(comp)
        if not exe:
            pytest.skip(f"{comp} not available")
        r = subprocess.run([exe, "-fsyntax-only", std, "-Wall", "-Werror", "-x", lang, hdr], capture_output=True,
                           text=True)
        assert r.returncode == 0, r.stderr
