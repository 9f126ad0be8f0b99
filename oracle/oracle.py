"""ctypes bindings of the CPU restatement (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_LIB_MT = os.path.join(_HERE, "liboracle_mt.so")

ORACLE_HEADER_NOTE = "CPU restatement of the reference path; test infrastructure only"


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB) or not os.path.exists(_LIB_MT):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


class Params(C.Structure):
    _fields_ = [("rho", C.c_double), ("mu", C.c_double), ("Delt", C.c_double),
                ("alfi", C.c_double), ("almi", C.c_double), ("gami", C.c_double),
                ("dtsfct", C.c_double), ("taucfct", C.c_double),
                ("flmpl", C.c_double), ("flmpr", C.c_double),
                ("ibksiz", C.c_int), ("lhs", C.c_int)]


_lib = None
_lib_mt = None


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = _bind(C.CDLL(_LIB))
    return _lib


def lib_mt(threads: int = 0):
    """The threaded build (OpenMP), with `threads` threads when > 0."""
    global _lib_mt
    if _lib_mt is None:
        build()
        _lib_mt = _bind(C.CDLL(_LIB_MT))
    if threads > 0:
        _lib_mt.cfo_threads(int(threads))
    return _lib_mt


_LIB_COUNT = os.path.join(_HERE, "liboracle_count.so")
_lib_count = None


def lib_count():
    """The counting build (flopcount.hpp: every +, -, *, / and sqrt of the
    restatement counted, the values unchanged)."""
    global _lib_count
    if _lib_count is None:
        if not os.path.exists(_LIB_COUNT):
            subprocess.run(["make", "-s", "-C", _HERE, _LIB_COUNT], check=True)
        L = _bind(C.CDLL(_LIB_COUNT))
        L.cfo_flop_count.restype = C.c_ulonglong
        L.cfo_flop_count.argtypes = [C.c_int]
        _lib_count = L
    return _lib_count


def count_interior_flops(ien, y, ac, x, colm, rowp, params):
    """Floating-point operations of one interior formation (flux pass, e3,
    scatter, fillsparseI) of the restatement on this mesh, counted by the
    counting build; returns (count, res, lhsK, lhsP)."""
    L = lib_count()
    ien = _f(ien, np.int32)
    y, ac, x = _f(y), _f(ac), _f(x)
    nshg, nnz = y.shape[0], rowp.shape[0]
    res = np.zeros((nshg, 4), order="F")
    lhsK = np.zeros((9, nnz), order="F")
    lhsP = np.zeros((4, nnz), order="F")
    L.cfo_flop_count(1)
    rc = L.cfo_elmgmr_interior(nshg, x.shape[0], ien.shape[0], ien, y, ac, x, None, _f(colm, np.int32),
                               _f(rowp, np.int32), nnz, res, _ptr(lhsK), _ptr(lhsP), None, C.byref(params))
    _check(rc, "elmgmr_interior (counting build)")
    return int(L.cfo_flop_count(1)), res, lhsK, lhsP


def _bind(L):
    if True:
        dp = np.ctypeslib.ndpointer(np.float64, flags="F_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="F_CONTIGUOUS")
        L.cfo_elmgmr_interior.argtypes = [C.c_int, C.c_int, C.c_int, ip, dp, dp, dp,
                                          C.c_void_p, ip, ip, C.c_int, dp, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.POINTER(Params)]
        L.cfo_elmgmr_interior_q.argtypes = [C.c_int, C.c_int, C.c_int, ip, dp, dp, dp, dp,
                                            C.c_void_p, ip, ip, C.c_int, dp, C.c_void_p,
                                            C.c_void_p, C.POINTER(Params)]
        L.cfo_e3_block.argtypes = [C.c_int, dp, dp, dp, dp, dp, dp, dp, C.POINTER(Params)]
        L.cfo_e3_block.restype = None
        L.cfo_e3q_block.argtypes = [C.c_int, dp, dp, dp, dp, C.POINTER(Params)]
        L.cfo_e3q_block.restype = None
        L.cfo_bc3res.argtypes = [C.c_int, ip, C.c_void_p, C.c_int, dp]
        L.cfo_elmgmr_interior_bc.argtypes = [C.c_int, C.c_int, C.c_int, ip, dp, dp, dp, ip, dp, C.c_int,
                                             ip, ip, C.c_int, dp, C.c_void_p, C.c_void_p, C.POINTER(Params)]
        L.cfo_bc3lhs_elem.argtypes = [C.c_int, ip, dp, ip, dp]
        L.cfo_bc3lhs_elem.restype = None
        L.cfo_boundary_lhs.argtypes = [C.c_int, C.c_int, ip, ip, dp, C.c_int, ip, ip, dp]
        L.cfo_genadj.argtypes = [C.c_int, C.c_int, ip, ip, ip, C.c_int, C.POINTER(C.c_int)]
        L.cfo_pack_val4.argtypes = [C.c_int, ip, ip, dp, dp, dp]
        L.cfo_pack_val4.restype = None
        L.cfo_lhs_create.argtypes = [C.c_int, C.c_int, ip, ip, C.c_int, C.POINTER(C.c_void_p)]
        L.cfo_lhs_free.argtypes = [C.c_void_p]
        L.cfo_lhs_free.restype = None
        L.cfo_bc_create.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, ip, C.c_void_p]
        L.cfo_ls_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]
        L.cfo_ls_free.argtypes = [C.c_void_p]
        L.cfo_ls_free.restype = None
        L.cfo_ls_solve.argtypes = [C.c_void_p, C.c_void_p, C.c_int, dp, dp, C.c_void_p, C.c_void_p]
        L.cfo_ls_stats.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
        L.cfo_ls_stats.restype = None
        L.cfo_time_params.argtypes = [C.c_double] + [C.POINTER(C.c_double)] * 3
        L.cfo_time_params.restype = None
        L.cfo_flops_per_element.restype = C.c_double
        L.cfo_boundary_res.argtypes = [C.c_int, C.c_int, C.c_int, ip, ip, dp, C.c_int, dp, dp,
                                       C.c_double, C.c_double, C.c_double, C.c_int, C.c_void_p,
                                       C.c_int, C.c_int, dp]
        L.cfo_boundary_res_flx.argtypes = [C.c_int, C.c_int, C.c_int, ip, ip, dp, C.c_int, dp, dp,
                                           C.c_double, C.c_double, C.c_double, C.c_int, C.c_void_p,
                                           C.c_int, C.c_int, dp, dp, C.c_int]
        L.cfo_boundary_wall.argtypes = [C.c_int, C.c_int, C.c_int, ip, ip, dp, C.c_int, dp, dp, dp, C.c_void_p,
                                        dp, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_void_p,
                                        C.c_int, C.c_int, C.POINTER(Wall), dp, C.c_void_p, C.c_void_p, C.c_int,
                                        C.c_void_p, C.c_void_p, C.c_void_p]
        L.cfo_plan_create.argtypes = [C.c_int, C.c_int, ip, C.c_int, C.POINTER(C.c_void_p)]
        L.cfo_plan_free.argtypes = [C.c_void_p]
        L.cfo_plan_free.restype = None
        L.cfo_elmgmr_interior_plan.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, ip, dp, dp, dp,
                                               C.c_void_p, C.c_void_p, C.c_int, ip, ip, C.c_int, dp, C.c_void_p,
                                               C.c_void_p, C.POINTER(Params)]
        L.cfo_threads.argtypes = [C.c_int]
    return L


def time_params(rhoinf: float):
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    lib().cfo_time_params(rhoinf, C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value  # almi, alfi, gami


def make_params(rho=1.06, mu=0.04, delt=1e-3, rhoinf=0.5, lhs=1, ibksiz=255):
    almi, alfi, gami = time_params(rhoinf)
    return Params(rho, mu, delt, alfi, almi, gami, 1.0, 1.0, 0.0, 0.0, ibksiz, lhs)


def _f(a, dt=np.float64):
    return np.asfortranarray(a, dtype=dt)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


def genadj(ien, nshg):
    """colm(nshg+1), rowp(nnz), 1-based (genadj.f90)."""
    ien = _f(ien, np.int32)
    if ien.size and (ien.min() < 1 or ien.max() > nshg):
        raise ValueError("genadj: ien entries must lie in 1..nshg")
    nel = ien.shape[0]
    cap = 32 * nshg
    colm = np.zeros(nshg + 1, np.int32)
    rowp = np.zeros(cap, np.int32)
    nnz = C.c_int()
    _check(lib().cfo_genadj(nshg, nel, ien, colm, rowp, cap, C.byref(nnz)), "genadj")
    return colm, rowp[: nnz.value].copy()


def elmgmr_interior(ien, y, ac, x, colm, rowp, params, ibc=None, want_q=False, q=None, bc=None):
    """Interior ElmGMR: returns res(nshg,4), lhsK(9,nnz), lhsP(4,nnz)[, qres].
    With bc = BC(nshg,ndofBC >= 6), partial velocity codes in ibc are
    projected out of the element tangents (bc3lhs)."""
    ien = _f(ien, np.int32)
    y, ac, x = _f(y), _f(ac), _f(x)
    nshg = y.shape[0]
    nnz = rowp.shape[0]
    res = np.zeros((nshg, 4), order="F")
    lhsK = np.zeros((9, nnz), order="F") if params.lhs == 1 else None
    lhsP = np.zeros((4, nnz), order="F") if params.lhs == 1 else None
    ibc_a = None if ibc is None else _f(ibc, np.int32)
    if bc is not None:
        if ibc is None or want_q or q is not None:
            raise ValueError("bc needs ibc, and neither want_q nor q")
        bc_a = _f(bc)
        rc = lib().cfo_elmgmr_interior_bc(nshg, x.shape[0], ien.shape[0], ien, y, ac, x, ibc_a, bc_a,
                                          bc_a.shape[1], _f(colm, np.int32), _f(rowp, np.int32), nnz, res,
                                          _ptr(lhsK), _ptr(lhsP), C.byref(params))
        _check(rc, "elmgmr_interior_bc")
        return res, lhsK, lhsP
    if q is not None:
        rc = lib().cfo_elmgmr_interior_q(nshg, x.shape[0], ien.shape[0], ien, y, ac, x, _f(q),
                                         _ptr(ibc_a), _f(colm, np.int32), _f(rowp, np.int32), nnz,
                                         res, _ptr(lhsK), _ptr(lhsP), C.byref(params))
        _check(rc, "elmgmr_interior_q")
        return res, lhsK, lhsP
    qres = np.zeros((nshg, 9), order="F") if want_q else None
    rc = lib().cfo_elmgmr_interior(nshg, x.shape[0], ien.shape[0], ien, y, ac, x, _ptr(ibc_a),
                                   _f(colm, np.int32), _f(rowp, np.int32), nnz, res,
                                   _ptr(lhsK), _ptr(lhsP), _ptr(qres), C.byref(params))
    _check(rc, "elmgmr_interior")
    return (res, lhsK, lhsP, qres) if want_q else (res, lhsK, lhsP)


class Plan:
    """cfo_plan over a mesh: the two-phase interior formation, serial
    (mt False) or on the threaded build's threads (mt True)."""

    def __init__(self, ien, nshg, ibksiz=255, mt=False):
        self.L = lib_mt() if mt else lib()
        self.ien = _f(ien, np.int32)
        self.nshg = int(nshg)
        h = C.c_void_p()
        _check(self.L.cfo_plan_create(self.nshg, self.ien.shape[0], self.ien, ibksiz, C.byref(h)), "plan_create")
        self.h = h

    def interior(self, y, ac, x, colm, rowp, params, ibc=None, bc=None):
        """res(nshg,4), lhsK(9,nnz), lhsP(4,nnz) as elmgmr_interior."""
        y, ac, x = _f(y), _f(ac), _f(x)
        nnz = rowp.shape[0]
        res = np.empty((self.nshg, 4), order="F")
        lhsK = np.empty((9, nnz), order="F") if params.lhs == 1 else None
        lhsP = np.empty((4, nnz), order="F") if params.lhs == 1 else None
        ibc_a = None if ibc is None else _f(ibc, np.int32)
        bc_a = None if bc is None else _f(bc)
        rc = self.L.cfo_elmgmr_interior_plan(self.h, self.nshg, x.shape[0], self.ien.shape[0], self.ien, y, ac,
                                             x, _ptr(ibc_a), _ptr(bc_a), 0 if bc is None else bc_a.shape[1],
                                             _f(colm, np.int32), _f(rowp, np.int32), nnz, res, _ptr(lhsK),
                                             _ptr(lhsP), C.byref(params))
        _check(rc, "elmgmr_interior_plan")
        return res, lhsK, lhsP

    def __del__(self):
        try:
            self.L.cfo_plan_free(self.h)
        except Exception:
            pass


def bc3lhs_elem(ibc, bc, nodes, xK):
    """bc3lhs on one element: xK(9,4,4) Fortran (K(m,a,b), a the row node),
    nodes 1-based; returns a new array."""
    out = _f(xK).copy(order="F")
    lib().cfo_bc3lhs_elem(int(np.asarray(ibc).shape[0]), _f(ibc, np.int32), _f(bc), _f(nodes, np.int32), out)
    return out


def boundary_lhs(ienb, ibc, bc, colm, rowp, lhsK):
    """lhsK(9,nnz) plus the boundary loop's tangent for rigid walls (bc3lhs's
    diagonal ones of partial-BC nodes, one per boundary element)."""
    out = _f(lhsK).copy(order="F")
    bc_a = _f(bc)
    _check(lib().cfo_boundary_lhs(bc_a.shape[0], int(np.asarray(ienb).shape[0]), _f(ienb, np.int32),
                                  _f(ibc, np.int32), bc_a, bc_a.shape[1], _f(colm, np.int32),
                                  _f(rowp, np.int32), out), "boundary_lhs")
    return out


def e3_block(yl, acl, xl, ql, params):
    """One e3 call on SoA block arrays (npro,4,4)... returns rl, xKebe, xGoC."""
    n = yl.shape[0]
    rl = np.zeros((n, 4, 4), order="F")
    K = np.zeros((n, 9, 4, 4), order="F")
    G = np.zeros((n, 4, 4, 4), order="F")
    lib().cfo_e3_block(n, _f(yl), _f(acl), _f(xl), _f(ql), rl, K, G, C.byref(params))
    return rl, K, G


def e3q_block(yl, xl, params):
    n = yl.shape[0]
    ql = np.zeros((n, 4, 9), order="F")
    rm = np.zeros((n, 4), order="F")
    lib().cfo_e3q_block(n, _f(yl), _f(xl), ql, rm, C.byref(params))
    return ql, rm


def bc3res(ibc, res, bc=None):
    res = _f(res).copy(order="F")
    bca = None if bc is None else _f(bc)
    _check(lib().cfo_bc3res(res.shape[0], _f(ibc, np.int32), _ptr(bca),
                            0 if bc is None else bc.shape[1], res), "bc3res")
    return res


def boundary_res(ienb, iBCB, BCB, y, x, res, rho=1.06, mu=0.04, influx=0.5, nqb=3, permit=None,
                 ibksiz=255):
    """res(nshg,4) plus the boundary elements' residual (AsBMFG/e3b/e3bvar):
    returns a new array.  ienb(nelb,4) 1-based with the face nodes first,
    iBCB(nelb,2), BCB(nelb,3,ndBCB), y(nshg,4) = (u1,u2,u3,p), x(numnp,3)."""
    res = _f(res).copy(order="F")
    ienb = _f(ienb, np.int32)
    iBCB = _f(iBCB, np.int32)
    BCB = _f(BCB)
    y = _f(y)
    x = _f(x)
    perm = None if permit is None else _f(permit, np.int32)
    nsrf = -1 if permit is None else perm.size - 1
    _check(lib().cfo_boundary_res(res.shape[0], x.shape[0], ienb.shape[0], ienb, iBCB, BCB, BCB.shape[2],
                                  y, x, rho, mu, influx, nqb, _ptr(perm), nsrf, ibksiz, res), "boundary_res")
    return res


def boundary_flux(ienb, iBCB, BCB, y, x, nflx, rho=1.06, mu=0.04, influx=0.5, nqb=3, permit=None,
                  ibksiz=255):
    """The boundary loop's surface integrals flxID(5, 0..nflx) by srfID
    (e3b.f90:216-247): area, outflow, force on the fluid (3)."""
    res = np.zeros((np.asarray(y).shape[0], 4), order="F")
    ienb = _f(ienb, np.int32)
    iBCB = _f(iBCB, np.int32)
    BCB = _f(BCB)
    perm = None if permit is None else np.ascontiguousarray(permit, dtype=np.int32)
    nsrf = -1 if permit is None else perm.size - 1
    flx = np.zeros((5, nflx + 1), order="F")
    _check(lib().cfo_boundary_res_flx(res.shape[0], np.asarray(x).shape[0], ienb.shape[0], ienb, iBCB, BCB,
                                      BCB.shape[2], _f(y), _f(x), rho, mu, influx, nqb, _ptr(perm), nsrf, ibksiz,
                                      res, flx, nflx), "boundary_flux")
    return flx


class Wall(C.Structure):
    """cfo_wall: the deformable-wall parameters and the time integration."""
    _fields_ = [("rhovw", C.c_double), ("thicknessvw", C.c_double), ("evw", C.c_double), ("rnuvw", C.c_double),
                ("rshearconstantvw", C.c_double), ("tissSuppStiffCoeff", C.c_double),
                ("tissSuppDampCoeff", C.c_double), ("tissSuppRingStiffCoeff", C.c_double),
                ("tissSuppRingDampCoeff", C.c_double), ("iwallmassfactor", C.c_int), ("iwallstiffactor", C.c_int),
                ("iwallsupp", C.c_int), ("iwalldamp", C.c_int), ("iringsupp", C.c_int), ("iringdamp", C.c_int),
                ("Delt", C.c_double), ("alfi", C.c_double), ("almi", C.c_double), ("gami", C.c_double),
                ("betai", C.c_double), ("flmpl", C.c_double)]


def make_wall(p=None, **kw):
    """cfo_wall with the reference basic case's wall values (testbin/mainTests/
    basic/solver.inp: rhovw 1.06, h 0.3, E 7e5, nu 0.5, kappa 0.833333,
    support 0.1, damping 0.01, ring 5e5 / 0.01, every term on) in the units
    of the formation tests (g, cm, s), the time integration of `p`
    (make_params) and betai = 1; keywords override any field."""
    p = p if p is not None else make_params()
    d = dict(rhovw=1.06, thicknessvw=0.3, evw=7.0e5, rnuvw=0.5, rshearconstantvw=0.833333,
             tissSuppStiffCoeff=0.1, tissSuppDampCoeff=0.01, tissSuppRingStiffCoeff=5e5, tissSuppRingDampCoeff=0.01,
             iwallmassfactor=1, iwallstiffactor=1, iwallsupp=1, iwalldamp=1, iringsupp=1, iringdamp=1,
             Delt=p.Delt, alfi=p.alfi, almi=p.almi, gami=p.gami, betai=1.0, flmpl=0.0)
    d.update(kw)
    return Wall(**d)


def boundary_wall(ienb, iBCB, BCB, y, ac, u, x, res, wall, uref=None, tag=None, rho=1.06, mu=0.04, influx=0.5,
                  nqb=3, permit=None, ibksiz=255, ibc=None, bc=None, colm=None, rowp=None, lhsK=None):
    """res(nshg,4) plus the boundary loop with deformable walls (e3b/e3bvar
    with ideformwall = 1); when lhsK(9,nnz) is given, each boundary element's
    xKebe (the wall tangent) through bc3lhs and fillsparseI is added to it.
    u, uref (nshg,3); tag(numnp) node tags (bit 0: ring)."""
    res = np.asfortranarray(res, dtype=np.float64)
    ienb = _f(ienb, np.int32)
    iBCB = _f(iBCB, np.int32)
    BCB = _f(BCB)
    perm = None if permit is None else np.ascontiguousarray(permit, dtype=np.int32)
    nsrf = -1 if permit is None else perm.size - 1
    nshg = res.shape[0]
    if lhsK is not None:
        ibc = np.ascontiguousarray(ibc if ibc is not None else np.zeros(nshg), dtype=np.int32)
        bc = _f(bc if bc is not None else np.zeros((nshg, 6)))
        colm, rowp = _f(colm, np.int32), _f(rowp, np.int32)
    tag_a = None if tag is None else np.ascontiguousarray(tag, dtype=np.int32)
    uref_a = None if uref is None else _f(uref)
    _check(lib().cfo_boundary_wall(nshg, np.asarray(x).shape[0], ienb.shape[0], ienb, iBCB, BCB, BCB.shape[2],
                                   _f(y), _f(ac), _f(u), _ptr(uref_a), _f(x), _ptr(tag_a), rho, mu, influx, nqb,
                                   _ptr(perm), nsrf, ibksiz, C.byref(wall), res,
                                   _ptr(ibc) if lhsK is not None else None,
                                   _ptr(bc) if lhsK is not None else None, bc.shape[1] if lhsK is not None else 0,
                                   _ptr(colm) if lhsK is not None else None,
                                   _ptr(rowp) if lhsK is not None else None, _ptr(lhsK)), "boundary_wall")
    return res


def pack_val4(colm, rowp, lhsK, lhsP, mt=False):
    nshg = colm.shape[0] - 1
    val = np.zeros((16, rowp.shape[0]), order="F")
    (lib_mt() if mt else lib()).cfo_pack_val4(nshg, _f(colm, np.int32), _f(rowp, np.int32), _f(lhsK), _f(lhsP), val)
    return val


LS_TYPE_CG, LS_TYPE_GMRES, LS_TYPE_NS = 1, 2, 3
BC_TYPE_Dir, BC_TYPE_Neu = 0, 1


class MemLS:
    """memLS_LHS_CREATE / BC_CREATE / LS_CREATE / SOLVE on the CPU."""

    def __init__(self, colm, rowp, nfaces, ls_type=LS_TYPE_NS, relTol=None, absTol=None,
                 maxItr=None, dimKry=None, relTolIn=None, absTolIn=None, maxItrIn=None, mt=False):
        self.L = L = lib_mt() if mt else lib()
        self.nNo = colm.shape[0] - 1
        self.nnz = rowp.shape[0]
        self._keep = []
        h = C.c_void_p()
        _check(L.cfo_lhs_create(self.nNo, self.nnz, _f(colm, np.int32), _f(rowp, np.int32), nfaces,
                                C.byref(h)), "lhs_create")
        self.lhs = h
        present = 0
        vals = dict(relTol=0.0, absTol=0.0, maxItr=0, dimKry=0)
        for bit, name, v in ((1, "relTol", relTol), (2, "absTol", absTol), (4, "maxItr", maxItr),
                             (8, "dimKry", dimKry)):
            if v is not None:
                present |= bit
                vals[name] = v
        arrs = []
        for bit, v, dt in ((16, relTolIn, np.float64), (32, absTolIn, np.float64), (64, maxItrIn, np.int32)):
            if v is not None:
                present |= bit
                a = np.asarray(v, dtype=dt)
                arrs.append(a)
            else:
                arrs.append(None)
        self._keep.extend(arrs)
        ls = C.c_void_p()
        _check(L.cfo_ls_create(ls_type, present, vals["relTol"], vals["absTol"], vals["maxItr"],
                               vals["dimKry"], _ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]),
                               C.byref(ls)), "ls_create")
        self.ls = ls

    def bc_create(self, faIn, gNodes, dof, bc_type, val=None):
        g = _f(gNodes, np.int32)
        v = None if val is None else _f(val)
        self._keep.append(v)
        _check(self.L.cfo_bc_create(self.lhs, faIn, g.shape[0], dof, bc_type, g, _ptr(v)), "bc_create")

    def solve(self, Ri, Val, incL=None, res=None):
        """Returns the solution (dof,nNo); Ri and Val are copied first."""
        Ri = _f(Ri).copy(order="F")
        Val = _f(Val).copy(order="F")
        il = None if incL is None else _f(incL, np.int32)
        rs = None if res is None else _f(res)
        _check(self.L.cfo_ls_solve(self.lhs, self.ls, Ri.shape[0], Ri, Val, _ptr(il), _ptr(rs)), "ls_solve")
        return Ri

    def stats(self, sub=0):
        out = (C.c_double * 8)()
        self.L.cfo_ls_stats(self.ls, sub, out)
        keys = ("iNorm", "fNorm", "itr", "suc", "dB", "callD", "Resm", "Resc")
        return dict(zip(keys, list(out)))

    def __del__(self):
        try:
            self.L.cfo_lhs_free(self.lhs)
            self.L.cfo_ls_free(self.ls)
        except Exception:
            pass


def flops_per_element():
    return lib().cfo_flops_per_element()


# ------------------------------------------------------------ step glue
# numpy restatements of the updates either side of the formation/solve in
# one nonlinear iteration (itrdrv.f90:1284-1528); y, ac (nshg,4) with
# columns (u1,u2,u3,p) as the solver arrays hold them.

def itr_predict(yold, acold, gami):
    """itrPredict, ipred = 1 (itrPC.f90:104-108): y = yold, ac = acold (gami-1)/gami."""
    return np.asfortranarray(yold.copy()), np.asfortranarray(acold * ((gami - 1.0) / gami))


def itr_yalpha(yold, acold, y, ac, alfi, almi):
    """itrYAlpha (itrPC.f90:326-334): the state at n+alpha; pressure at n+1,
    the pressure column of acAlpha zero."""
    ya = np.asfortranarray(y.copy())
    aca = np.zeros_like(ac, order="F")
    ya[:, :3] = yold[:, :3] + alfi * (y[:, :3] - yold[:, :3])
    aca[:, :3] = acold[:, :3] + almi * (ac[:, :3] - acold[:, :3])
    return ya, aca


def itr_correct(y, ac, sol4, gami, alfi, delt):
    """itrCorrect (itrPC.f90:176-181) with solinc(i,1:4) = Ri(1:4,i)."""
    y = np.asfortranarray(y.copy())
    ac = np.asfortranarray(ac.copy())
    d = np.asarray(sol4).T
    y[:, :3] += (gami * delt) * d[:, :3]
    ac[:, :3] += d[:, :3]
    y[:, 3] += (gami * alfi * delt) * d[:, 3]
    return y, ac


def itr_bc(ibc, bc, y):
    """itrBC (itrbc.f90:20-98) without periodic nodes: code 7 sets the
    velocity to BC(:,3:5); codes 1-6 set the constrained components from
    BC(:,3:6) and the free ones; bit 2 sets the pressure to BC(:,1)."""
    y = np.asfortranarray(y.copy())
    code = (ibc >> 3) & 7
    u1, u2, u3 = y[:, 0].copy(), y[:, 1].copy(), y[:, 2].copy()
    b = [bc[:, k] for k in range(min(bc.shape[1], 6))] + [np.zeros(y.shape[0])] * max(0, 6 - bc.shape[1])
    for c, comps in ((1, {0: b[2] - b[3] * u2 - b[4] * u3}),
                     (2, {1: b[2] - b[3] * u1 - b[4] * u3}),
                     (3, {0: b[2] - b[3] * u3, 1: b[4] - b[5] * u3}),
                     (4, {2: b[2] - b[3] * u1 - b[4] * u2}),
                     (5, {0: b[2] - b[3] * u2, 2: b[4] - b[5] * u2}),
                     (6, {1: b[2] - b[3] * u1, 2: b[4] - b[5] * u1}),
                     (7, {0: b[2], 1: b[3], 2: b[4]})):
        sel = code == c
        for k, v in comps.items():
            y[sel, k] = v[sel]
    p = ((ibc >> 2) & 1) == 1
    y[p, 3] = bc[p, 0]
    return y


def elmpvsq_coef(res, y, glob, nabi, c1, c2):
    """ElmpvsQ's term of a coupled surface with implicit coefficients
    (elmgmr.f90:1210-1228 for RCR, the same form for every 0D model):
    res(glob,1:3) += -(c1 Q + c2) NABI."""
    res = np.asfortranarray(res.copy())
    q = float(np.sum(nabi * y[glob, :3]))
    res[glob, :3] += (-(c1 * q + c2)) * nabi
    return res, q


def elmpvsq_resistance(res, y, glob, nabi, R):
    """ElmpvsQ's resistance term (elmgmr.f90:1152-1176): Q = sum NABI.u over
    the face's nodes, res(glob,1:3) += -R Q NABI."""
    res = np.asfortranarray(res.copy())
    q = float(np.sum(nabi * y[glob, :3]))
    res[glob, :3] += (-q * R) * nabi
    return res, q
