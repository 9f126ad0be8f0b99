"""crimson_form_boundary_flux_noalloc (the surface integrals staged in the
library's fixed device array instead of a buffer the handle allocates)
against crimson_form_boundary_flux, both called through the C ABI: bitwise
equal on two calls in a row, the second on a smaller mesh with fewer
surfaces and a column count reaching past them (the array still holds the
first call's larger sums there); a srfID past the array's limit is refused.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfa():
    import crimsonflowsolver_amd as cfa
    if cfa.device_count() < 1:
        pytest.fail("no GPU visible: the HIP path must run here")
    return cfa


def _case(cfa, grid, nsurf, seed):
    """A formed cylinder whose boundary elements carry srfIDs 1..nsurf (some
    negative, which count as their absolute value)."""
    from crimsonflowsolver_amd import mesh as M
    m = M.cylinder_mesh(*grid)
    colm, rowp = cfa.genadj(m.ien, m.numnp)
    y, ac = M.poiseuille_state(m, seed=seed)
    ienb, iBCB, BCB = M.cylinder_boundary(m)
    rng = np.random.default_rng(seed)
    srf = rng.integers(1, nsurf + 1, size=ienb.shape[0]) * rng.choice([-1, 1], size=ienb.shape[0])
    iBCB = np.asfortranarray(np.stack([iBCB[:, 0], srf], axis=1).astype(np.int32))
    form = cfa.ElmGMRInterior(m.ien, colm, rowp, m.numnp, params=cfa.form_params())
    form.set_boundary(ienb, iBCB, BCB)
    form(y, ac, m.x)
    return form, (y, ac, m.x)


def _flux(cfa, form, nflx, noalloc):
    out = np.full((5, nflx + 1), np.nan, order="F")
    fn = cfa.lib().crimson_form_boundary_flux_noalloc if noalloc else cfa.lib().crimson_form_boundary_flux
    rc = fn(form.h, int(nflx), out.ctypes.data_as(C.c_void_p))
    return rc, out


def test_noalloc_flux_equals_allocating_entry_point(cfa):
    big, _ = _case(cfa, (8, 8, 16), 9, seed=3)
    print("case: 8x8x16 cylinder, 9 surfaces, nflx 9", flush=True)
    rc0, want = _flux(cfa, big, 9, noalloc=False)
    rc1, got = _flux(cfa, big, 9, noalloc=True)
    assert rc0 == 0 and rc1 == 0
    assert np.all(want[:, 1:] != 0.0) and np.all(want[:, 0] == 0.0)
    assert np.array_equal(got, want)

    # second call: a smaller mesh with 4 surfaces, asked for 9 columns
    small, _ = _case(cfa, (6, 6, 8), 4, seed=5)
    print("case: 6x6x8 cylinder, 4 surfaces, nflx 9", flush=True)
    rc0, want = _flux(cfa, small, 9, noalloc=False)
    rc1, got = _flux(cfa, small, 9, noalloc=True)
    assert rc0 == 0 and rc1 == 0
    assert np.all(want[:, 5:] == 0.0) and np.all(want[:, 1:5] != 0.0)
    assert np.array_equal(got, want)
    # fewer columns than surfaces
    _, want = _flux(cfa, small, 2, noalloc=False)
    _, got = _flux(cfa, small, 2, noalloc=True)
    assert np.array_equal(got, want)


def test_noalloc_flux_refuses_srfid_past_its_limit(cfa):
    lim = cfa.FLUX_NOALLOC_MAX_SURFACES
    form, state = _case(cfa, (6, 6, 8), lim, seed=11)
    ib = form._bnd[1].copy(order="F")
    ib[0, 1] = lim  # |srfID| = limit: one column past the array
    form.set_boundary(form._bnd[0], ib, form._bnd[2])
    form(*state)
    print(f"case: srfID {lim} (limit {lim})", flush=True)
    rc, _ = _flux(cfa, form, 4, noalloc=True)
    assert rc == 1  # CRIMSON_EARG, before any launch
    assert "srfID" in cfa.last_error()
    rc, out = _flux(cfa, form, 4, noalloc=False)  # the allocating entry point still serves it
    assert rc == 0 and np.all(np.isfinite(out))
